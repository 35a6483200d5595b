/*
 * lifdiff_oracle.c -- CPU restatement of the reference's differentiable LIF:
 * forward with a tape, checkpoint-segment replay, and the reverse sweep.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP adjoint in
 * fly-brain_amd/: only tests/ and bench.py's CPU baseline leg may load it.
 *
 * What it restates (reference paths are into Lulzx/fly-brain):
 *   - inScale from sizeLog, and the delivered weight    src/lifdiff.js:158-164, 278-285
 *   - forward: deliver, driven pre-pass in set order,   src/lifdiff.js:219-353
 *     fused integrate / decay / threshold, spike lists
 *   - checkpoints and segment replay with the arena     src/lifdiff.js:358-464
 *   - backward: reversed decays, mode dispatch,         src/lifdiff.js:494-686
 *     surrogate (fast sigmoid, or the logistic's own
 *     derivative in soft mode), reversed arrivals,
 *     adjoint clip, truncation, non-finite sanitising
 *   - the drive as an input: a schedule of drive rates  src/lifdiff.js:135-144, 228-234,
 *     per epoch for the neurons of driveIdx, applied     338, 406-413, 453, 518-527, 565-605
 *     by setDriveOne at each epoch start (the onEpoch
 *     hook), driveSoft (the drive's expectation as a
 *     graded spike), the drive gradient per epoch, and
 *     per-epoch spike weights (dLdSpikePerEpoch)
 *   - exact32: the forward pass in lif.c's binary32 arithmetic,     src/lifdiff.js:105-109, 173-190,
 *     with writeGraph's baked weights and signs                      231-240, 263-274, 317-326
 *   - a caller-given inScale, and a per-neuron surrogateBeta         src/lifdiff.js:89-95, 115-118, 160
 *   - forward's spikeTrain: listed neurons forced like driven        src/lifdiff.js:219, 243, 298-307,
 *     spikes after the draws, kept with the tape for the replay      432-438
 * The reference evaluates in JavaScript doubles and stores into Float32Arrays:
 * every local below is a double and every array store rounds to binary32, in
 * the reference's evaluation order.  Under exact32 every forward intermediate
 * is binary32 instead, one rounding per C operation in lif.c's order.  One
 * line differs from lifdiff.js there: the depression of res.  lifdiff.js
 * writes `res[pre] -= dU * res[pre]`, one binary64 rounding before the store,
 * where lif.c rounds the product to binary32 first; at depU = 0.15 the two
 * disagree on about 9% of updates.  exact32 exists to be lif.c bit for bit,
 * so this file takes lif.c's rounding (the two agree whenever depU is 0, the
 * setting the reference audits it in).
 *
 * Two additions the reference does not have, restated from the same
 * derivation:
 *   - gEdge[j] = dL/d edgeGain[j], the per-edge multiplier lifdiff.js applies
 *     where counts enter the delivery (src/lifdiff.js:126-129, 282).  Each
 *     arrival adds lg[q] * counts[j] * inScale[q] * eff (* inhGain); summed
 *     over a row this is the arrival's gLogGain term.
 *   - a filtered-rate readout loss: r_t = a r_{t-1} + (1-a) s_t / (dt/1000),
 *     a = exp(-dt/tau), L = mean over steps and readouts of (r_t - target)^2.
 *     Its spike adjoint is the reverse filter
 *     lam_t = 2 (r_t - target) / (T R) + a lam_{t+1},  dL/ds_t = (1-a)/(dt/1000) lam_t,
 *     added to the flat dLdSpike term where the reference adds dLdSpike.
 *
 * Pinning: tests/test_lifdiff_oracle.py checks the gradients against central
 * finite differences of this file's own soft forward (the method and floor of
 * scripts/grad_check.mjs:92-99), the row / column invariants of gEdge, and the
 * hard forward's spike raster against the lif.c restatement where the two
 * models coincide.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "lifdiff_oracle.h"
#include "lif_oracle.h"

#define FORCED_BIT 0x80000000u

typedef struct {
    float *v, *gE, *gI, *adapt, *res, *refr;
    uint32_t rng;
    int32_t *drv, nDrv;   /* the driven set at the checkpoint (it changes under a drive schedule) */
} ckpt_t;

#define DRIVE_MAX 0.999   /* src/lifdiff.js:72 */

struct fdo_net {
    int32_t N;
    int64_t E;
    uint32_t *indptr, *indices;
    float *counts, *edgeGain, *sign;
    float *in[FDO_NINPUTS];
    uint8_t *thrMask, *biasMask;
    int32_t inGiven[FDO_NINPUTS];   /* FDO_INSCALE_IN / FDO_BETA set */
    float *W32, *signW32;           /* exact32: writeGraph's W and sign (src/lifdiff.js:177-190) */
    int32_t *driven, nDriven;
    fdo_params p;
    int32_t nslots, K;
    /* state */
    float *st[FDO_NSTATE];
    float *scored;
    uint32_t rng;
    /* loss */
    float *dLdSpike;
    int32_t lossFrom;
    int32_t R;
    int32_t *readout, *readIdx;
    float *target;
    double tauR;
    /* tape */
    int32_t steps;
    int64_t *off;
    int32_t *idx;
    float *amp;
    int64_t cap;
    float *readS;
    double *rate;
    ckpt_t *ck;
    int32_t nck;
    /* drive schedule: rates per epoch of `ep` steps for the neurons drvIdx[0..nDI) */
    int32_t nDI, ep, nEp, driveSoft;
    int32_t *drvIdx, *drvSlot;
    float *drvSched;
    float *gDrive;                /* [nEp][nDI] after backward */
    float *dPer;                  /* [nEpPer][N] per-epoch spike weights, or NULL */
    int32_t nEpPer;
    double lossPer;               /* the per-epoch weighted loss of the last forward */
    /* the drive and driven set of a run in progress (forward or replay) */
    float *wdrive;
    int32_t *wdrv, nW;
    uint8_t *inW;
    /* the spike train for the next forward pass, and the one the tape ran with */
    int32_t trainSteps, tapeTrainSteps;
    int64_t *trainPtr, *tapeTrainPtr;
    int32_t *trainIdx, *tapeTrainIdx;
};

static void *xcalloc(size_t n, size_t s) { return calloc(n ? n : 1, s); }

static uint32_t xs(uint32_t *s) { return fbo_xorshift32(s); }
/* (s >>> 8) * 2^-24, as LIFDiff._rand returns it (src/lifdiff.js:213) */
static double rand01(uint32_t *s) { return (double)(xs(s) >> 8) * (1.0 / 16777216.0); }

static double js_round(double x) {
    double f = floor(x);
    return (x - f >= 0.5) ? f + 1.0 : f;
}

fdo_net *fdo_create(int32_t N, int64_t E, const uint32_t *indptr, const uint32_t *indices,
                    const float *counts, const float *sign, const fdo_params *p) {
    fdo_net *n = (fdo_net *)calloc(1, sizeof *n);
    n->N = N;
    n->E = E;
    n->indptr = (uint32_t *)xcalloc(N + 1, 4);
    memcpy(n->indptr, indptr, 4 * (size_t)(N + 1));
    n->indices = (uint32_t *)xcalloc(E, 4);
    memcpy(n->indices, indices, 4 * (size_t)E);
    n->counts = (float *)xcalloc(E, 4);
    memcpy(n->counts, counts, 4 * (size_t)E);
    n->edgeGain = NULL;
    n->sign = (float *)xcalloc(N, 4);
    memcpy(n->sign, sign, 4 * (size_t)N);
    for (int k = 0; k < FDO_NINPUTS; k++) n->in[k] = (float *)xcalloc(N, 4);
    for (int32_t i = 0; i < N; i++) n->in[FDO_OUTSCALE][i] = 1.f;
    n->thrMask = (uint8_t *)xcalloc(N, 1);
    n->biasMask = (uint8_t *)xcalloc(N, 1);
    n->driven = (int32_t *)xcalloc(N, 4);
    n->p = *p;
    double r = js_round(p->delay / p->dt);
    n->nslots = (int32_t)(r < 1 ? 1 : r) + 1;
    n->K = p->checkpoint > 0 ? p->checkpoint : 16;
    for (int k = 0; k < FDO_NSTATE; k++) n->st[k] = (float *)xcalloc(N, 4);
    n->scored = (float *)xcalloc(N, 4);
    n->dLdSpike = (float *)xcalloc(N, 4);
    n->readIdx = (int32_t *)xcalloc(N, 4);
    for (int32_t i = 0; i < N; i++) n->readIdx[i] = -1;
    n->drvSlot = (int32_t *)xcalloc(N, 4);
    for (int32_t i = 0; i < N; i++) n->drvSlot[i] = -1;
    n->ep = 1;
    n->wdrive = (float *)xcalloc(N, 4);
    n->wdrv = (int32_t *)xcalloc(N, 4);
    n->inW = (uint8_t *)xcalloc(N, 1);
    return n;
}

static void free_tape(fdo_net *n) {
    free(n->off); free(n->idx); free(n->amp); free(n->readS); free(n->rate);
    n->off = NULL; n->idx = NULL; n->amp = NULL; n->readS = NULL; n->rate = NULL;
    for (int32_t c = 0; c < n->nck; c++) {
        free(n->ck[c].v); free(n->ck[c].gE); free(n->ck[c].gI);
        free(n->ck[c].adapt); free(n->ck[c].res); free(n->ck[c].refr); free(n->ck[c].drv);
    }
    free(n->ck);
    n->ck = NULL;
    n->nck = 0;
    free(n->tapeTrainPtr); free(n->tapeTrainIdx);
    n->tapeTrainPtr = NULL; n->tapeTrainIdx = NULL;
    n->tapeTrainSteps = 0;
}

void fdo_destroy(fdo_net *n) {
    if (!n) return;
    free(n->W32); free(n->signW32);
    free_tape(n);
    free(n->indptr); free(n->indices); free(n->counts); free(n->edgeGain); free(n->sign);
    for (int k = 0; k < FDO_NINPUTS; k++) free(n->in[k]);
    for (int k = 0; k < FDO_NSTATE; k++) free(n->st[k]);
    free(n->thrMask); free(n->biasMask); free(n->driven); free(n->scored); free(n->dLdSpike);
    free(n->readout); free(n->readIdx); free(n->target);
    free(n->drvIdx); free(n->drvSlot); free(n->drvSched); free(n->gDrive); free(n->dPer);
    free(n->wdrive); free(n->wdrv); free(n->inW);
    free(n->trainPtr); free(n->trainIdx);
    free(n);
}

/* forward's spikeTrain (src/lifdiff.js:219, 298-307): step t < steps forces
 * idx[ptr[t] .. ptr[t+1]) in that order; steps = 0 clears.  The next forward
 * pass copies it into its tape, whose replay applies it (src/lifdiff.js:243,
 * 432-438). */
void fdo_set_spike_train(fdo_net *n, int32_t steps, const int64_t *ptr, const int32_t *idx) {
    free(n->trainPtr); free(n->trainIdx);
    n->trainPtr = NULL; n->trainIdx = NULL;
    n->trainSteps = 0;
    if (steps <= 0) return;
    n->trainPtr = (int64_t *)malloc(8 * ((size_t)steps + 1));
    memcpy(n->trainPtr, ptr, 8 * ((size_t)steps + 1));
    n->trainIdx = (int32_t *)xcalloc((size_t)ptr[steps], 4);
    if (ptr[steps]) memcpy(n->trainIdx, idx, 4 * (size_t)ptr[steps]);
    n->trainSteps = steps;
}

/* driveIdx / driveEpoch with the rates the onEpoch hook would set
 * (src/lifdiff.js:135-144, 228-234): at the start of epoch e the rate of
 * drvIdx[k] becomes sched[e*nD + k], through setDriveOne's set semantics.
 * driveSoft replaces the Poisson drive by its expectation (src/lifdiff.js:338). */
void fdo_set_drive_schedule(fdo_net *n, const int32_t *idx, int32_t nD, int32_t epochSteps, const float *sched,
                            int32_t nEp, int32_t driveSoft) {
    for (int32_t k = 0; k < n->nDI; k++) n->drvSlot[n->drvIdx[k]] = -1;
    free(n->drvIdx); free(n->drvSched); free(n->gDrive);
    n->drvIdx = NULL; n->drvSched = NULL; n->gDrive = NULL;
    n->nDI = idx && sched ? nD : 0;
    n->nEp = n->nDI ? nEp : 0;
    n->ep = epochSteps > 0 ? epochSteps : 1;
    n->driveSoft = driveSoft;
    if (!n->nDI) return;
    n->drvIdx = (int32_t *)xcalloc(nD, 4);
    memcpy(n->drvIdx, idx, 4 * (size_t)nD);
    for (int32_t k = 0; k < nD; k++) n->drvSlot[idx[k]] = k;
    n->drvSched = (float *)xcalloc((size_t)nEp * nD, 4);
    memcpy(n->drvSched, sched, 4 * (size_t)nEp * nD);
    n->gDrive = (float *)xcalloc((size_t)nEp * nD, 4);
}

/* dLdSpikePerEpoch (src/lifdiff.js:519-522): the weight of a spike of neuron i
 * at a step of epoch e is dPer[e*N + i] instead of dLdSpike[i]. */
void fdo_set_loss_per_epoch(fdo_net *n, const float *dPer, int32_t nEp, int32_t epochSteps) {
    free(n->dPer);
    n->dPer = NULL;
    n->nEpPer = 0;
    if (!dPer) return;
    n->dPer = (float *)xcalloc((size_t)nEp * n->N, 4);
    memcpy(n->dPer, dPer, 4 * (size_t)nEp * n->N);
    n->nEpPer = nEp;
    n->ep = epochSteps > 0 ? epochSteps : 1;
}

void fdo_drive_grad(const fdo_net *n, float *gDrive) {
    if (n->gDrive && gDrive) memcpy(gDrive, n->gDrive, 4 * (size_t)n->nEp * n->nDI);
}

/* setDriveOne on the run's working drive and driven set (src/lifdiff.js:205-207):
 * an added neuron goes to the end, a removed one leaves its place */
/* setDriveOne (src/lifdiff.js:207): the neuron joins the driven set when
 * the rate is positive.  A rate of -0.0 stands for a positive binary64 rate
 * that the binary32 drive array holds as zero: it joins with a zero drive. */
static void drive_set_one(fdo_net *n, int32_t i, float rate) {
    n->wdrive[i] = rate;
    if (rate > 0 || (rate == 0 && signbit(rate))) {
        if (!n->inW[i]) { n->inW[i] = 1; n->wdrv[n->nW++] = i; }
    } else if (n->inW[i]) {
        n->inW[i] = 0;
        int32_t w = 0;
        for (int32_t q = 0; q < n->nW; q++) if (n->wdrv[q] != i) n->wdrv[w++] = n->wdrv[q];
        n->nW = w;
    }
}

/* the schedule's rates for epoch e (the last epoch's rates hold beyond it) */
static void apply_epoch(fdo_net *n, int32_t e) {
    if (!n->nDI) return;
    if (e >= n->nEp) e = n->nEp - 1;
    for (int32_t k = 0; k < n->nDI; k++) drive_set_one(n, n->drvIdx[k], n->drvSched[(int64_t)e * n->nDI + k]);
}

/* the working drive and driven set at the start of a run: the inputs */
static void start_drive(fdo_net *n) {
    memcpy(n->wdrive, n->in[FDO_DRIVE], 4 * (size_t)n->N);
    memset(n->inW, 0, (size_t)n->N);
    n->nW = 0;
    for (int32_t q = 0; q < n->nDriven; q++) { n->wdrv[n->nW++] = n->driven[q]; n->inW[n->driven[q]] = 1; }
}

void fdo_set_input(fdo_net *n, int32_t which, const float *values) {
    if (which < 0 || which >= FDO_NINPUTS) return;
    n->inGiven[which] = values != NULL;
    if (values) memcpy(n->in[which], values, 4 * (size_t)n->N);
}

void fdo_set_edge_gain(fdo_net *n, const float *eg) {
    if (!eg) { free(n->edgeGain); n->edgeGain = NULL; return; }
    if (!n->edgeGain) n->edgeGain = (float *)xcalloc(n->E, 4);
    memcpy(n->edgeGain, eg, 4 * (size_t)n->E);
}

void fdo_set_masks(fdo_net *n, const uint8_t *thrMask, const uint8_t *biasMask) {
    if (thrMask) memcpy(n->thrMask, thrMask, (size_t)n->N);
    if (biasMask) memcpy(n->biasMask, biasMask, (size_t)n->N);
}

/* drivenSet in insertion order (src/lifdiff.js:203-207) */
void fdo_set_driven(fdo_net *n, const int32_t *list, int32_t count) {
    memcpy(n->driven, list, 4 * (size_t)count);
    n->nDriven = count;
}

void fdo_set_loss(fdo_net *n, const float *dLdSpike, int32_t lossFrom) {
    memcpy(n->dLdSpike, dLdSpike, 4 * (size_t)n->N);
    n->lossFrom = lossFrom;
}

void fdo_set_readout(fdo_net *n, const int32_t *idx, int32_t R, const float *targetHz, double tauMs) {
    free(n->readout); free(n->target);
    for (int32_t i = 0; i < n->N; i++) n->readIdx[i] = -1;
    n->R = R;
    n->readout = (int32_t *)xcalloc(R, 4);
    n->target = (float *)xcalloc(R, 4);
    memcpy(n->readout, idx, 4 * (size_t)R);
    memcpy(n->target, targetHz, 4 * (size_t)R);
    for (int32_t k = 0; k < R; k++) n->readIdx[idx[k]] = k;
    n->tauR = tauMs;
}

float *fdo_state(fdo_net *n, int32_t which) { return n->st[which]; }
uint32_t fdo_rng(const fdo_net *n) { return n->rng; }

/* ------------------------------------------------------------------ constants */

typedef struct {
    double dE, dA, kRec, kM, dtS, cE, cI;
} kconsts;

static kconsts consts_of(const fdo_params *p) {
    kconsts k;
    k.dE = exp(-p->dt / p->tauSyn);
    k.dA = exp(-p->dt / p->adaptTau);
    k.kRec = p->dt / p->depTau;
    k.kM = p->dt / p->tauM;
    k.dtS = p->dt / 1000.0;
    k.cE = 1.0 / (p->eExc - p->vRest);
    k.cI = 1.0 / (p->vRest - p->eInh);
    return k;
}

/* inScale = min(boostCap, max(1/maxSizeScale, exp(-sizeAlpha * sizeLog)))
 * (src/lifdiff.js:161-164) */
static void prepare(fdo_net *n) {
    const fdo_params *p = &n->p;
    float *in = n->st[FDO_INSCALE];
    if (n->inGiven[FDO_INSCALE_IN]) {
        memcpy(in, n->in[FDO_INSCALE_IN], 4 * (size_t)n->N);   /* taken verbatim */
    } else {
        for (int32_t i = 0; i < n->N; i++) {
            double s = exp(-p->sizeAlpha * (double)n->in[FDO_SIZELOG][i]);
            double lo = 1.0 / p->maxSizeScale;
            double m = s > lo ? s : lo;
            in[i] = (float)(p->boostCap < m ? p->boostCap : m);
        }
    }
    /* exact32 bakes writeGraph's arrays (src/lifdiff.js:177-190): W[k] =
     * f32(count * edgeGain * inScale[q] * ig) with ig = (sign < 0 ? inhGain :
     * 1) * outScale[j], and sign[j] = f32(sign * wSyn), each product in f64 */
    free(n->W32); free(n->signW32);
    n->W32 = NULL; n->signW32 = NULL;
    if (!p->exact32) return;
    n->W32 = (float *)xcalloc(n->E, 4);
    n->signW32 = (float *)xcalloc(n->N, 4);
    for (int32_t j = 0; j < n->N; j++) {
        const double sg = n->sign[j];
        n->signW32[j] = (float)(sg * p->wSyn);
        const double ig = (sg < 0 ? p->inhGain : 1.0) * (double)n->in[FDO_OUTSCALE][j];
        for (uint32_t k = n->indptr[j]; k < n->indptr[j + 1]; k++)
            n->W32[k] = (float)((double)n->counts[k] * (n->edgeGain ? (double)n->edgeGain[k] : 1.0) *
                                (double)in[n->indices[k]] * ig);
    }
}

/* a neuron's surrogateBeta: the per-neuron array when given, else the scalar */
static double beta_of(const fdo_net *n, int32_t i) {
    return n->inGiven[FDO_BETA] ? (double)n->in[FDO_BETA][i] : n->p.surrogateBeta;
}

static void reset_state(fdo_net *n) {
    const fdo_params *p = &n->p;
    for (int32_t i = 0; i < n->N; i++) {
        n->st[FDO_V][i] = (float)p->vRest;
        n->st[FDO_GE][i] = 0.f; n->st[FDO_GI][i] = 0.f; n->st[FDO_ADAPT][i] = 0.f;
        n->st[FDO_RES][i] = 1.f; n->st[FDO_REFR][i] = 0.f; n->st[FDO_SPIKECOUNT][i] = 0.f;
        n->scored[i] = 0.f;
    }
    prepare(n);
}

/* eff = sign * wSyn * outScale * exp(logGain) * res * amp (src/lifdiff.js:278) */
static double eff_of(const fdo_net *n, int32_t pre, double res, double amp) {
    return (double)n->sign[pre] * n->p.wSyn * (double)n->in[FDO_OUTSCALE][pre] *
           exp((double)n->in[FDO_LOGGAIN][pre]) * res * amp;
}

/* the arrivals of one step, in list order (src/lifdiff.js:259-286).  resAt,
 * when given, records res[pre] as read. */
static void deliver(fdo_net *n, int64_t a0, int64_t a1, float *resAt) {
    float *gE = n->st[FDO_GE], *gI = n->st[FDO_GI], *res = n->st[FDO_RES];
    const float *inS = n->st[FDO_INSCALE];
    const float *eg = n->edgeGain;
    for (int64_t k = a0; k < a1; k++) {
        const int32_t pre = (int32_t)((uint32_t)n->idx[k] & ~FORCED_BIT);
        const double amp = n->amp[k];
        if (n->sign[pre] == 0.f) continue;
        if (resAt) resAt[k - a0] = res[pre];
        const double eff = eff_of(n, pre, res[pre], amp);
        res[pre] = (float)((double)res[pre] - n->p.depU * (double)res[pre]);
        const uint32_t a = n->indptr[pre], b = n->indptr[pre + 1];
        if (eff > 0) {
            for (uint32_t j = a; j < b; j++) {
                const uint32_t q = n->indices[j];
                gE[q] = (float)((double)gE[q] + (double)n->counts[j] * (eg ? (double)eg[j] : 1.0) * (double)inS[q] * eff);
            }
        } else {
            const double ig = n->p.inhGain;
            for (uint32_t j = a; j < b; j++) {
                const uint32_t q = n->indices[j];
                gI[q] = (float)((double)gI[q] + (double)n->counts[j] * (eg ? (double)eg[j] : 1.0) * (double)inS[q] * ig * eff);
            }
        }
    }
}

/* exact32's delivery (src/lifdiff.js:263-274): lif.c's factorisation -- s =
 * f32(sign * res), one binary32 product per edge into the conductance, a zero
 * s skipped before the depression.  The depression rounds as lif.c does (see
 * the header). */
static void deliver_x32(fdo_net *n, int64_t a0, int64_t a1) {
    float *gE = n->st[FDO_GE], *gI = n->st[FDO_GI], *res = n->st[FDO_RES];
    const float dU = (float)n->p.depU;
    for (int64_t k = a0; k < a1; k++) {
        const int32_t pre = (int32_t)((uint32_t)n->idx[k] & ~FORCED_BIT);
        const float amp = n->amp[k];
        float s = n->signW32[pre] * res[pre];
        if (amp != 1.f) s = s * amp;
        if (s == 0.f) continue;
        res[pre] = res[pre] - dU * res[pre];
        const uint32_t a = n->indptr[pre], b = n->indptr[pre + 1];
        float *g = s > 0 ? gE : gI;
        for (uint32_t j = a; j < b; j++) {
            const float dlt = n->W32[j] * s;
            if (dlt != 0.f) g[n->indices[j]] += dlt;
        }
    }
}

static void push_spike(fdo_net *n, uint32_t idx, float amp) {
    /* appends to the list of the step in progress, n->steps */
    if (n->off[n->steps + 1] >= n->cap) {
        n->cap = n->cap * 2 + 1024;
        n->idx = (int32_t *)realloc(n->idx, 4 * (size_t)n->cap);
        n->amp = (float *)realloc(n->amp, 4 * (size_t)n->cap);
    }
    const int64_t k = n->off[n->steps + 1]++;
    n->idx[k] = (int32_t)idx;
    n->amp[k] = amp;
}

static void save_ckpt(fdo_net *n, int32_t c) {
    ckpt_t *ck = &n->ck[c];
    const size_t sz = 4 * (size_t)n->N;
    if (!ck->v) {
        ck->v = (float *)malloc(sz); ck->gE = (float *)malloc(sz); ck->gI = (float *)malloc(sz);
        ck->adapt = (float *)malloc(sz); ck->res = (float *)malloc(sz); ck->refr = (float *)malloc(sz);
    }
    memcpy(ck->v, n->st[FDO_V], sz); memcpy(ck->gE, n->st[FDO_GE], sz); memcpy(ck->gI, n->st[FDO_GI], sz);
    memcpy(ck->adapt, n->st[FDO_ADAPT], sz); memcpy(ck->res, n->st[FDO_RES], sz);
    memcpy(ck->refr, n->st[FDO_REFR], sz);
    ck->rng = n->rng;
    free(ck->drv);
    ck->drv = (int32_t *)xcalloc(n->nW, 4);
    memcpy(ck->drv, n->wdrv, 4 * (size_t)n->nW);
    ck->nDrv = n->nW;
}

static void restore_ckpt(fdo_net *n, int32_t c) {
    const ckpt_t *ck = &n->ck[c];
    const size_t sz = 4 * (size_t)n->N;
    memcpy(n->st[FDO_V], ck->v, sz); memcpy(n->st[FDO_GE], ck->gE, sz); memcpy(n->st[FDO_GI], ck->gI, sz);
    memcpy(n->st[FDO_ADAPT], ck->adapt, sz); memcpy(n->st[FDO_RES], ck->res, sz);
    memcpy(n->st[FDO_REFR], ck->refr, sz);
    n->rng = ck->rng;
    memcpy(n->wdrive, n->in[FDO_DRIVE], sz);
    memset(n->inW, 0, (size_t)n->N);
    n->nW = ck->nDrv;
    memcpy(n->wdrv, ck->drv, 4 * (size_t)ck->nDrv);
    for (int32_t q = 0; q < n->nW; q++) n->inW[n->wdrv[q]] = 1;
}

/* filtered rates of the readouts and the loss terms */
static double readout_loss(fdo_net *n) {
    const int32_t T = n->steps, R = n->R;
    if (!R || !T) return 0.0;
    const double a = exp(-n->p.dt / n->tauR), dtS = n->p.dt / 1000.0;
    double L = 0.0;
    for (int32_t k = 0; k < R; k++) {
        double r = 0.0;
        for (int32_t t = 0; t < T; t++) {
            r = a * r + (1.0 - a) * (double)n->readS[(int64_t)t * R + k] / dtS;
            n->rate[(int64_t)t * R + k] = r;
            const double d = r - (double)n->target[k];
            L += d * d;
        }
    }
    return L / ((double)T * (double)R);
}

double fdo_forward(fdo_net *n, int32_t steps, double seed, int32_t record) {
    const fdo_params *p = &n->p;
    const int32_t N = n->N, D = n->nslots - 1, R = n->R;
    const kconsts k = consts_of(p);
    free_tape(n);
    n->steps = 0;
    n->off = (int64_t *)xcalloc((size_t)steps + 1, 8);
    n->cap = 1024;
    n->idx = (int32_t *)malloc(4 * (size_t)n->cap);
    n->amp = (float *)malloc(4 * (size_t)n->cap);
    n->readS = (float *)xcalloc((size_t)steps * (R ? R : 1), 4);
    n->rate = (double *)xcalloc((size_t)steps * (R ? R : 1), 8);
    if (record) {
        n->nck = (steps + n->K - 1) / n->K + 1;
        n->ck = (ckpt_t *)calloc((size_t)n->nck, sizeof(ckpt_t));
    }
    n->rng = fbo_seed_hash(seed);
    reset_state(n);
    start_drive(n);
    n->lossPer = 0.0;
    if (n->trainSteps) {   /* the tape keeps the train it ran with */
        const int64_t nI = n->trainPtr[n->trainSteps];
        n->tapeTrainPtr = (int64_t *)malloc(8 * ((size_t)n->trainSteps + 1));
        memcpy(n->tapeTrainPtr, n->trainPtr, 8 * ((size_t)n->trainSteps + 1));
        n->tapeTrainIdx = (int32_t *)xcalloc((size_t)nI, 4);
        if (nI) memcpy(n->tapeTrainIdx, n->trainIdx, 4 * (size_t)nI);
        n->tapeTrainSteps = n->trainSteps;
    }
    const int ex = p->exact32 != 0;
    if (ex) {
        /* lif.c sums vThresh + adapt + thr only: the additive field must be in
         * thrOffset (src/lifdiff.js:178-180) */
        for (int32_t i = 0; i < N; i++)
            if (n->thrMask[i]) return NAN;
    }
    /* exact32: every scalar as lif.c reads it from its binary32 header */
    const float xvR = (float)p->vRest, xeE = (float)p->eExc, xeI = (float)p->eInh, xvTh = (float)p->vThresh,
                xvRs = (float)p->vReset, xtRf = (float)p->tRef, xaI = (float)p->adaptInc, xdT = (float)p->dt,
                xlB = (float)p->laminaBias, xdE = (float)k.dE, xdA = (float)k.dA, xkR = (float)k.kRec,
                xkM = (float)k.kM, xdtS = (float)k.dtS, xcE = (float)k.cE, xcI = (float)k.cI;
    float *v = n->st[FDO_V], *gE = n->st[FDO_GE], *gI = n->st[FDO_GI], *adapt = n->st[FDO_ADAPT];
    float *res = n->st[FDO_RES], *refr = n->st[FDO_REFR], *sc = n->st[FDO_SPIKECOUNT];
    const float *drive = n->wdrive, *thrOff = n->in[FDO_THROFFSET];
    for (int32_t t = 0; t < steps; t++) {
        if (record && t % n->K == 0) save_ckpt(n, t / n->K);
        if (n->nDI && t % n->ep == 0) apply_epoch(n, t / n->ep);   /* the onEpoch hook */
        n->off[t + 1] = n->off[t];
        if (t >= D) {
            if (ex) deliver_x32(n, n->off[t - D], n->off[t - D + 1]);
            else deliver(n, n->off[t - D], n->off[t - D + 1], NULL);
        }
        n->steps = t;
        const int scoredStep = t >= n->lossFrom;
        const float *wPer = n->dPer ? n->dPer + (int64_t)(t / n->ep) * N : NULL;
        /* a forced spike (src/lifdiff.js:293-296, 303-305) */
#define FORCE_SPIKE(i)                                                                   \
    do {                                                                                 \
        refr[i] = ex ? xtRf + xdT : (float)(p->tRef + p->dt);                            \
        push_spike(n, (uint32_t)(i) | FORCED_BIT, 1.f);                                  \
        sc[i] = (float)((double)sc[i] + 1.0);                                            \
        if (scoredStep) n->scored[i] = (float)((double)n->scored[i] + 1.0);              \
        if (scoredStep && wPer) n->lossPer += (double)wPer[i];                           \
        adapt[i] = ex ? adapt[i] + xaI : (float)((double)adapt[i] + p->adaptInc);        \
        if (R && n->readIdx[i] >= 0) n->readS[(int64_t)t * R + n->readIdx[i]] = 1.f;     \
    } while (0)
        /* driven pre-pass in set order (src/lifdiff.js:290-297) */
        for (int32_t q = 0; q < n->nW && !n->driveSoft; q++) {
            const int32_t i = n->wdrv[q];
            if (refr[i] > 0) continue;
            if (rand01(&n->rng) < (ex ? (double)(drive[i] * xdtS) : (double)drive[i] * k.dtS)) FORCE_SPIKE(i);
        }
        /* the spike train: listed neurons in list order, no draw (src/lifdiff.js:298-307) */
        if (t < n->trainSteps)
            for (int64_t kk = n->trainPtr[t]; kk < n->trainPtr[t + 1]; kk++) {
                const int32_t i = n->trainIdx[kk];
                if (refr[i] > 0) continue;
                FORCE_SPIKE(i);
            }
#undef FORCE_SPIKE
        /* integrate, decay, threshold (src/lifdiff.js:313-345) */
        for (int32_t i = 0; i < N; i++) {
            double vi = v[i], s = 0.0;
            const int wasRefr = refr[i] > 0;
            if (wasRefr) {
                refr[i] = ex ? refr[i] - xdT : (float)((double)refr[i] - p->dt);
                vi = ex ? (double)xvRs : p->vReset;
            } else if (ex) {
                /* lif.c's COBA expression, one binary32 rounding per operation */
                const float vf = v[i], bias = n->biasMask[i] ? xlB : 0.f;
                float acc = xvR - vf;
                acc = acc + (gE[i] * (xeE - vf)) * xcE;
                acc = acc + (gI[i] * (vf - xeI)) * xcI;
                vi = (double)(vf + (acc + bias) * xkM);
            } else {
                const double bias = n->biasMask[i] ? p->laminaBias : 0.0;
                vi += (p->vRest - vi + (double)gE[i] * (p->eExc - vi) * k.cE + (double)gI[i] * (vi - p->eInh) * k.cI + bias) * k.kM;
            }
            if (ex) {
                gE[i] = gE[i] * xdE;
                gI[i] = gI[i] * xdE;
                adapt[i] = adapt[i] * xdA;
                res[i] = res[i] + (1.f - res[i]) * xkR;
            } else {
                gE[i] = (float)((double)gE[i] * k.dE);
                gI[i] = (float)((double)gI[i] * k.dE);
                adapt[i] = (float)((double)adapt[i] * k.dA);
                res[i] = (float)((double)res[i] + (1.0 - (double)res[i]) * k.kRec);
            }
            if (!wasRefr) {
                const double thr = ex ? (double)((xvTh + adapt[i]) + thrOff[i])
                                      : p->vThresh + (double)adapt[i] + (n->thrMask[i] ? p->kcThreshold : 0.0) + (double)thrOff[i];
                s = p->soft ? 1.0 / (1.0 + exp(-(vi - thr) / beta_of(n, i))) : (vi >= thr ? 1.0 : 0.0);
                /* the drive as an expectation (src/lifdiff.js:338) */
                if (n->driveSoft && drive[i] > 0) {
                    const double sd = fmin(DRIVE_MAX, (double)drive[i] * k.dtS);
                    s = sd + (1.0 - sd) * s;
                }
                if (s > 0) {
                    vi = s * (ex ? (double)xvRs : p->vReset) + (1.0 - s) * vi;
                    if (!p->soft) refr[i] = ex ? xtRf : (float)p->tRef;
                    adapt[i] = (float)((double)adapt[i] + s * (ex ? (double)xaI : p->adaptInc));
                }
                sc[i] = (float)((double)sc[i] + s);
                if (scoredStep) n->scored[i] = (float)((double)n->scored[i] + s);
                if (scoredStep && wPer) n->lossPer += (double)wPer[i] * s;
                if (R && n->readIdx[i] >= 0) n->readS[(int64_t)t * R + n->readIdx[i]] = (float)s;
                if (s > (p->soft ? 1e-4 : 0.5)) push_spike(n, (uint32_t)i, (float)s);
            }
            v[i] = (float)vi;
        }
    }
    n->steps = steps;
    double L = 0.0;
    if (n->dPer) L = n->lossPer;
    else for (int32_t i = 0; i < N; i++) L += (double)n->dLdSpike[i] * (double)n->scored[i];
    return L + readout_loss(n);
}

int64_t fdo_spikes(const fdo_net *n, int64_t *off, int32_t *idx, float *amp, int64_t cap) {
    const int64_t tot = n->off ? n->off[n->steps] : 0;
    if (off) memcpy(off, n->off, 8 * ((size_t)n->steps + 1));
    const int64_t m = tot < cap ? tot : cap;
    if (idx && m > 0) memcpy(idx, n->idx, 4 * (size_t)m);
    if (amp && m > 0) memcpy(amp, n->amp, 4 * (size_t)m);
    return tot;
}

int64_t fdo_readout(const fdo_net *n, float *s, double *rate) {
    const int64_t m = (int64_t)n->steps * n->R;
    if (s && m) memcpy(s, n->readS, 4 * (size_t)m);
    if (rate && m) memcpy(rate, n->rate, 8 * (size_t)m);
    return m;
}

/* ------------------------------------------------------------------ replay */

typedef struct {
    int32_t L;
    float *vPre, *gE, *gI, *adapt, *sp, *resAt;
    uint8_t *mode;
} arena_t;

/* Re-simulates [from, to) from the checkpoint at `from`, recording per step the
 * membrane potential, conductances (after delivery, before decay), decayed
 * adapt, spike amplitude, mode (0 integrated, 1 refractory, 2 driven spike,
 * 3 integrated with the soft drive) and res at each arrival
 * (src/lifdiff.js:390-464). */
static void replay(fdo_net *n, int32_t from, int32_t to, arena_t *A) {
    const fdo_params *p = &n->p;
    const int32_t N = n->N, D = n->nslots - 1;
    const kconsts k = consts_of(p);
    restore_ckpt(n, from / n->K);
    float *v = n->st[FDO_V], *gE = n->st[FDO_GE], *gI = n->st[FDO_GI], *adapt = n->st[FDO_ADAPT];
    float *res = n->st[FDO_RES], *refr = n->st[FDO_REFR];
    const float *drive = n->wdrive, *thrOff = n->in[FDO_THROFFSET];
    uint8_t *forced = (uint8_t *)xcalloc(N, 1);
    int64_t rOff = 0;
    for (int32_t t = from; t < to; t++) {
        const int64_t off = (int64_t)(t - from) * N;
        /* the rates the forward used: a segment need not start on an epoch
         * (src/lifdiff.js:406-413); reapplying them is idempotent */
        if (n->nDI && (t == from || t % n->ep == 0)) apply_epoch(n, t / n->ep);
        if (t >= D) {
            deliver(n, n->off[t - D], n->off[t - D + 1], A->resAt + rOff);
            rOff += n->off[t - D + 1] - n->off[t - D];
        }
        memset(forced, 0, (size_t)N);
        for (int32_t q = 0; q < n->nW && !n->driveSoft; q++) {
            const int32_t i = n->wdrv[q];
            if (refr[i] > 0) continue;
            if (rand01(&n->rng) < (double)drive[i] * k.dtS) {
                refr[i] = (float)(p->tRef + p->dt);
                adapt[i] = (float)((double)adapt[i] + p->adaptInc);
                forced[i] = 1;
            }
        }
        /* the tape's spike train, as forward applied it (src/lifdiff.js:432-438) */
        if (t < n->tapeTrainSteps)
            for (int64_t kk = n->tapeTrainPtr[t]; kk < n->tapeTrainPtr[t + 1]; kk++) {
                const int32_t i = n->tapeTrainIdx[kk];
                if (refr[i] > 0) continue;
                refr[i] = (float)(p->tRef + p->dt);
                adapt[i] = (float)((double)adapt[i] + p->adaptInc);
                forced[i] = 1;
            }
        for (int32_t i = 0; i < N; i++) {
            double vi = v[i], sv = 0.0;
            uint8_t md = 0;
            A->vPre[off + i] = v[i]; A->gE[off + i] = gE[i]; A->gI[off + i] = gI[i];
            const int wasRefr = refr[i] > 0;
            if (wasRefr) {
                refr[i] = (float)((double)refr[i] - p->dt);
                vi = p->vReset;
                if (forced[i]) { md = 2; sv = 1.0; } else md = 1;
            } else {
                const double bias = n->biasMask[i] ? p->laminaBias : 0.0;
                vi += (p->vRest - vi + (double)gE[i] * (p->eExc - vi) * k.cE + (double)gI[i] * (vi - p->eInh) * k.cI + bias) * k.kM;
            }
            gE[i] = (float)((double)gE[i] * k.dE);
            gI[i] = (float)((double)gI[i] * k.dE);
            adapt[i] = (float)((double)adapt[i] * k.dA);
            res[i] = (float)((double)res[i] + (1.0 - (double)res[i]) * k.kRec);
            A->adapt[off + i] = adapt[i];
            if (!wasRefr) {
                const double thr = p->vThresh + (double)adapt[i] + (n->thrMask[i] ? p->kcThreshold : 0.0) + (double)thrOff[i];
                sv = p->soft ? 1.0 / (1.0 + exp(-(vi - thr) / beta_of(n, i))) : (vi >= thr ? 1.0 : 0.0);
                if (n->driveSoft && drive[i] > 0) {
                    const double sd = fmin(DRIVE_MAX, (double)drive[i] * k.dtS);
                    sv = sd + (1.0 - sd) * sv;
                    md = 3;
                }
                if (sv > 0) {
                    vi = sv * p->vReset + (1.0 - sv) * vi;
                    if (!p->soft) refr[i] = (float)p->tRef;
                    adapt[i] = (float)((double)adapt[i] + sv * p->adaptInc);
                }
            }
            A->mode[off + i] = md;
            A->sp[off + i] = (float)sv;
            v[i] = (float)vi;
        }
    }
    free(forced);
}

/* ------------------------------------------------------------------ backward */

static float clipf(double x, double c) {
    if (c > 0) { if (x > c) x = c; else if (x < -c) x = -c; }
    return (float)x;
}

void fdo_backward(fdo_net *n, int32_t truncate, double *params, float *gLogGain, float *gEdge, float *gInScale,
                  int32_t *clamped) {
    const fdo_params *p = &n->p;
    const int32_t N = n->N, T = n->steps, R = n->R, S = n->nslots, D = S - 1;
    const kconsts k = consts_of(p);
    const double aClip = p->adjClip;
    const float *inS = n->st[FDO_INSCALE], *thrOff = n->in[FDO_THROFFSET];
    const float *eg = n->edgeGain;
    float *lv = (float *)xcalloc(N, 4), *lgE = (float *)xcalloc(N, 4), *lgI = (float *)xcalloc(N, 4);
    float *lad = (float *)xcalloc(N, 4), *lres = (float *)xcalloc(N, 4);
    float *ring = (float *)xcalloc((size_t)S * N, 4);
    memset(gLogGain, 0, 4 * (size_t)N);
    memset(gEdge, 0, 4 * (size_t)n->E);
    memset(gInScale, 0, 4 * (size_t)N);
    if (n->gDrive) memset(n->gDrive, 0, 4 * (size_t)n->nEp * n->nDI);
    double g[FDO_NPARAMS] = {0};
    double sAdaptInc = 0, sVThresh = 0, sKc = 0, sLamin = 0, sEInh = 0;
    /* the readout loss's spike adjoint, per step and readout */
    double *dLds = NULL;
    if (R) {
        const double a = exp(-p->dt / n->tauR), dtS = p->dt / 1000.0;
        dLds = (double *)xcalloc((size_t)T * R, 8);
        for (int32_t q = 0; q < R; q++) {
            double lam = 0.0;
            for (int32_t t = T - 1; t >= 0; t--) {
                lam = 2.0 * (n->rate[(int64_t)t * R + q] - (double)n->target[q]) / ((double)T * (double)R) + a * lam;
                dLds[(int64_t)t * R + q] = (1.0 - a) / dtS * lam;
            }
        }
    }
    arena_t A;
    A.L = n->K;
    A.vPre = (float *)xcalloc((size_t)A.L * N, 4); A.gE = (float *)xcalloc((size_t)A.L * N, 4);
    A.gI = (float *)xcalloc((size_t)A.L * N, 4); A.adapt = (float *)xcalloc((size_t)A.L * N, 4);
    A.sp = (float *)xcalloc((size_t)A.L * N, 4); A.mode = (uint8_t *)xcalloc((size_t)A.L * N, 1);
    int64_t resCap = 1;
    for (int32_t from = 0; from < T; from += n->K) {
        const int32_t to = from + n->K < T ? from + n->K : T;
        const int64_t c = (to > D ? n->off[to - D] : 0) - (from > D ? n->off[from - D] : 0);
        if (c > resCap) resCap = c;
    }
    A.resAt = (float *)xcalloc((size_t)resCap, 4);

    const int32_t nseg = (T + n->K - 1) / n->K;
    for (int32_t sgi = nseg - 1; sgi >= 0; sgi--) {
        const int32_t from = sgi * n->K, to = from + n->K < T ? from + n->K : T;
        replay(n, from, to, &A);
        for (int32_t t = to - 1; t >= from; t--) {
            const int64_t off = (int64_t)(t - from) * N;
            const int scoredStep = t >= n->lossFrom;
            const int32_t e = t / n->ep;
            const float *wPer = n->dPer ? n->dPer + (int64_t)e * N : NULL;
            float *gDr = n->gDrive ? n->gDrive + (int64_t)(e < n->nEp ? e : n->nEp - 1) * n->nDI : NULL;
            if (truncate > 0 && (t + 1) % truncate == 0) {
                memset(lv, 0, 4 * (size_t)N); memset(lgE, 0, 4 * (size_t)N); memset(lgI, 0, 4 * (size_t)N);
                memset(lad, 0, 4 * (size_t)N); memset(lres, 0, 4 * (size_t)N);
                memset(ring, 0, 4 * (size_t)S * N);
            }
            float *ls = ring + (int64_t)(t % S) * N;
            /* reversed decays and integrate-and-spike (src/lifdiff.js:555-617) */
            for (int32_t i = 0; i < N; i++) {
                const uint8_t md = A.mode[off + i];
                const double lpA = lad[i];
                const double lE = (double)lgE[i] * k.dE, lI = (double)lgI[i] * k.dE, lA = lpA * k.dA;
                const double lR = (double)lres[i] * (1.0 - k.kRec);
                lgE[i] = (float)lE; lgI[i] = (float)lI; lad[i] = (float)lA; lres[i] = (float)lR;
                double dl = scoredStep ? (wPer ? (double)wPer[i] : (double)n->dLdSpike[i]) : 0.0;
                if (R && n->readIdx[i] >= 0) dl += dLds[(int64_t)t * R + n->readIdx[i]];
                const int32_t slot = n->drvSlot[i];
                if (md == 1) { lv[i] = 0.f; continue; }
                if (md == 2) {
                    /* the drive fired: only the adaptation increment and the
                     * drive's own gradient remain (src/lifdiff.js:565-574) */
                    sAdaptInc += lA;
                    if (slot >= 0 && gDr)
                        gDr[slot] = (float)((double)gDr[slot] + (lA * p->adaptInc + (double)ls[i] + dl) * k.dtS);
                    lv[i] = 0.f; ls[i] = 0.f;
                    continue;
                }
                const double vv = A.vPre[off + i], gEv = A.gE[off + i], gIv = A.gI[off + i];
                const double bias = n->biasMask[i] ? p->laminaBias : 0.0;
                const double u = vv + (p->vRest - vv + gEv * (p->eExc - vv) * k.cE + gIv * (vv - p->eInh) * k.cI + bias) * k.kM;
                const double thr = p->vThresh + (double)A.adapt[off + i] + (n->thrMask[i] ? p->kcThreshold : 0.0) + (double)thrOff[i];
                const double d = u - thr;
                const double s = A.sp[off + i];
                /* with the soft drive the amplitude is sd + (1 - sd) st; the
                 * surrogate is taken at st (src/lifdiff.js:580-585) */
                double sd = 0.0, st = s;
                if (md == 3) {
                    const double rate = slot >= 0 ? (double)n->drvSched[(int64_t)(e < n->nEp ? e : n->nEp - 1) * n->nDI + slot]
                                                  : (double)n->in[FDO_DRIVE][i];
                    sd = fmin(DRIVE_MAX, rate * k.dtS);
                    st = (s - sd) / (1.0 - sd);
                }
                const double beta = beta_of(n, i);
                double sg;
                if (p->soft) sg = st * (1.0 - st) / beta;
                else { const double q = 1.0 + fabs(d) / beta; sg = 1.0 / (beta * q * q); }
                const double lspike = (double)ls[i] + dl;
                const double lvNew = lv[i];
                const double lsTot = (p->vReset - u) * lvNew + lpA * p->adaptInc + lspike;
                const double lt = lsTot * (1.0 - sd) * sg;
                const double lu = lvNew * (1.0 - s) + lt;
                const double lthr = -lt;
                /* and through sd: ds/d(sd) = 1 - st, d(sd)/d(rate) = dt/1000 */
                if (slot >= 0 && gDr) gDr[slot] = (float)((double)gDr[slot] + lsTot * (1.0 - st) * k.dtS);
                sAdaptInc += lpA * s;
                sVThresh += lthr;
                if (n->thrMask[i]) sKc += lthr;
                lad[i] = (float)(lA + lthr * k.dA);
                lv[i] = (float)(lu * (1.0 + k.kM * (-1.0 - gEv * k.cE + gIv * k.cI)));
                lgE[i] = (float)(lE + lu * k.kM * (p->eExc - vv) * k.cE);
                lgI[i] = (float)(lI + lu * k.kM * (vv - p->eInh) * k.cI);
                if (bias != 0.0) sLamin += lu * k.kM;
                sEInh += lu * k.kM * gIv * k.cI * (-1.0 + (vv - p->eInh) * k.cI);
                ls[i] = 0.f;
            }
            if (aClip > 0)
                for (int32_t i = 0; i < N; i++) {
                    lv[i] = clipf(lv[i], aClip); lgE[i] = clipf(lgE[i], aClip); lgI[i] = clipf(lgI[i], aClip);
                    lad[i] = clipf(lad[i], aClip); lres[i] = clipf(lres[i], aClip);
                }
            /* reversed delivery (src/lifdiff.js:635-666), plus gEdge */
            if (t < D) continue;
            const int64_t a0 = n->off[t - D], a1 = n->off[t - D + 1];
            /* this step's first entry in the segment's resAt record */
            const int64_t ro = n->off[t - D] - n->off[(from > D ? from : D) - D];
            const int32_t emit = t - D;
            float *target = emit >= 0 ? ring + (int64_t)(emit % S) * N : NULL;
            for (int64_t kk = a0; kk < a1; kk++) {
                const int32_t pre = (int32_t)((uint32_t)n->idx[kk] & ~FORCED_BIT);
                const double sgn = n->sign[pre], amp = n->amp[kk];
                if (sgn == 0.0) continue;
                const double resAtv = A.resAt[ro + (kk - a0)];
                const double gain = exp((double)n->in[FDO_LOGGAIN][pre]) * (double)n->in[FDO_OUTSCALE][pre];
                const double base = sgn * p->wSyn * gain * resAtv;
                const double eff = base * amp;
                const uint32_t a = n->indptr[pre], b = n->indptr[pre + 1];
                double lRaw = 0.0;
                if (eff > 0) {
                    for (uint32_t j = a; j < b; j++) {
                        const uint32_t q = n->indices[j];
                        const double cw = (double)n->counts[j] * (eg ? (double)eg[j] : 1.0);
                        lRaw += (double)lgE[q] * cw * (double)inS[q];
                        gInScale[q] = (float)((double)gInScale[q] + (double)lgE[q] * cw * eff);
                        gEdge[j] = (float)((double)gEdge[j] + (double)lgE[q] * (double)n->counts[j] * (double)inS[q] * eff);
                    }
                } else {
                    const double ig = p->inhGain;
                    for (uint32_t j = a; j < b; j++) {
                        const uint32_t q = n->indices[j];
                        const double cw = (double)n->counts[j] * (eg ? (double)eg[j] : 1.0);
                        lRaw += (double)lgI[q] * cw * (double)inS[q] * ig;
                        gInScale[q] = (float)((double)gInScale[q] + (double)lgI[q] * cw * ig * eff);
                        g[FDO_G_INHGAIN] += (double)lgI[q] * cw * (double)inS[q] * eff;
                        gEdge[j] = (float)((double)gEdge[j] + (double)lgI[q] * (double)n->counts[j] * (double)inS[q] * ig * eff);
                    }
                }
                g[FDO_G_WSYN] += lRaw * sgn * gain * resAtv * amp;
                gLogGain[pre] = (float)((double)gLogGain[pre] + lRaw * eff);
                if (target) target[pre] = clipf((double)target[pre] + lRaw * base, aClip);
                const double lresAfter = lres[pre];
                g[FDO_G_DEPU] += -lresAfter * resAtv;
                lres[pre] = clipf(lresAfter * (1.0 - p->depU) + lRaw * sgn * p->wSyn * gain * amp, aClip);
            }
        }
    }
    g[FDO_G_ADAPTINC] += sAdaptInc; g[FDO_G_VTHRESH] += sVThresh; g[FDO_G_KC] += sKc;
    g[FDO_G_LAMINA] += sLamin; g[FDO_G_EINH] += sEInh;
    /* sizeAlpha reaches the loss through inScale (src/lifdiff.js:670-674) */
    for (int32_t i = 0; i < N; i++) {
        const double raw = exp(-p->sizeAlpha * (double)n->in[FDO_SIZELOG][i]);
        if (raw < p->boostCap && raw > 1.0 / p->maxSizeScale)
            g[FDO_G_SIZEALPHA] += (double)gInScale[i] * (-(double)n->in[FDO_SIZELOG][i]) * raw;
    }
    int32_t cl[3] = {0, 0, 0};
    for (int q = 0; q < FDO_NPARAMS; q++) if (!isfinite(g[q])) { g[q] = 0; cl[0]++; }
    for (int32_t i = 0; i < N; i++) if (!isfinite(gLogGain[i])) { gLogGain[i] = 0.f; cl[1]++; }
    for (int64_t j = 0; j < n->E; j++) if (!isfinite(gEdge[j])) { gEdge[j] = 0.f; cl[2]++; }
    if (n->gDrive)
        for (int64_t q = 0; q < (int64_t)n->nEp * n->nDI; q++) if (!isfinite(n->gDrive[q])) n->gDrive[q] = 0.f;
    memcpy(params, g, sizeof g);
    if (clamped) memcpy(clamped, cl, sizeof cl);
    free(lv); free(lgE); free(lgI); free(lad); free(lres); free(ring); free(dLds);
    free(A.vPre); free(A.gE); free(A.gI); free(A.adapt); free(A.sp); free(A.mode); free(A.resAt);
}
