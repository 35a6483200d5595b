/*
 * lif_oracle.c -- CPU restatement of the reference's connectome LIF step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP engine in
 * fly-brain_amd/: only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it.  The product path never calls it.
 *
 * What it restates, phase by phase (reference paths are into Lulzx/fly-brain):
 *   - brain header constants and seed hash      src/lifwasm.js:39,52,55-68
 *   - reset / driven-list semantics             src/lifwasm.js:70-78
 *   - lif_step: deliver, background, driven,     src/wasm/lif.c:29-80
 *     membrane (CUBA / COBA), threshold + ring
 *   - xorshift32 with the (13, 17, 5) triple     src/wasm/lif.c:27
 *
 * Arithmetic is IEEE binary32 with no contraction: build with
 * -ffp-contract=off (the reference runs as wasm32, which has no fused
 * multiply-add).  Every expression below keeps the association the C
 * reference evaluates, which is what makes rasters bit-identical.
 *
 * Pinning: tests/test_oracle.py checks this file against golden vectors
 * produced by the reference lif.c itself (oracle/_ref, built by
 * oracle/Makefile from /root/reference), plus closed forms for the membrane,
 * decay and delay line, and brute-force xorshift streams.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#include "lif_oracle.h"

/* ------------------------------------------------------------------ helpers */

uint32_t fbo_xorshift32(uint32_t *state) {
    uint32_t x = *state;
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    *state = x;
    return x;
}

/* ECMAScript Math.round for finite positive and negative doubles: nearest
 * integer, ties toward +infinity. */
static double js_round(double x) {
    double f = floor(x);
    return (x - f >= 0.5) ? f + 1.0 : f;
}

/* (seed * 2654435761) >>> 0 || 1, evaluated the way a JS engine does:
 * an f64 product, then ToUint32. */
uint32_t fbo_seed_hash(double seed) {
    double prod = seed * 2654435761.0;
    double t = prod < 0 ? -floor(-prod) : floor(prod);
    double m = fmod(t, 4294967296.0);
    if (m < 0) m += 4294967296.0;
    uint32_t s = (uint32_t)m;
    return s ? s : 1u;
}

int32_t fbo_nslots(double delay, double dt) {
    double r = js_round(delay / dt);
    if (r < 1.0) r = 1.0;
    return (int32_t)r + 1;
}

void fbo_constants(const fbo_params *p, int32_t N, fbo_consts *c) {
    /* every derived constant is computed in f64 and rounded once to f32 */
    c->dt = (float)p->dt;
    c->vRest = (float)p->vRest;
    c->vThresh = (float)p->vThresh;
    c->vReset = (float)p->vReset;
    c->tRef = (float)p->tRef;
    c->adaptInc = (float)p->adaptInc;
    c->depU = (float)p->depU;
    c->dE = (float)exp(-p->dt / p->tauSyn);
    c->dTr = (float)exp(-p->dt / p->traceTau);
    c->dA = (float)exp(-p->dt / p->adaptTau);
    c->kRec = (float)(p->dt / p->depTau);
    c->kM = (float)(p->dt / p->tauM);
    c->dtS = (float)(p->dt / 1000.0);
    c->eExc = (float)p->eExc;
    c->eInh = (float)p->eInh;
    c->cE = (float)(1.0 / (p->eExc - p->vRest));
    c->cI = (float)(1.0 / (p->vRest - p->eInh));
    c->bgEvents = (float)((double)N * p->bgRate * p->dt / 1000.0);
    c->bgAmp = (float)p->bgAmp;
    c->preInh = (float)p->preInh;
    c->coba = p->coba != 0.0;
}

/* ------------------------------------------------------------------ brain */

struct fbo_brain {
    int32_t N, nslots;
    int64_t E;
    fbo_consts k;
    /* graph, copied so the caller's buffers may go away */
    uint32_t *indptr, *indices;
    float *W, *sign;
    /* per-neuron state */
    float *arr[FBO_NARRAYS];
    uint32_t *spikeCount;
    /* delay line: nslots lists of up to N neuron ids */
    int32_t *ring, *ringCount;
    int32_t *driven, nDriven;
    /* scalars */
    int32_t head, lastSlot, nFired;
    uint32_t rng;
    float bgAcc;
};

#define V(b) ((b)->arr[FBO_V])
#define GE(b) ((b)->arr[FBO_GE])
#define GI(b) ((b)->arr[FBO_GI])
#define REFR(b) ((b)->arr[FBO_REFR])
#define TRACE(b) ((b)->arr[FBO_TRACE])
#define ADAPT(b) ((b)->arr[FBO_ADAPT])
#define RES(b) ((b)->arr[FBO_RES])
#define BIAS(b) ((b)->arr[FBO_BIAS])
#define THR(b) ((b)->arr[FBO_THR])
#define DRIVE(b) ((b)->arr[FBO_DRIVE])

void fbo_reset(fbo_brain *b) {
    /* lifwasm.js:78 -- RNG, head and the driven list survive a reset */
    int32_t N = b->N;
    for (int32_t i = 0; i < N; i++) {
        V(b)[i] = b->k.vRest;
        GE(b)[i] = 0.f; GI(b)[i] = 0.f; REFR(b)[i] = 0.f;
        TRACE(b)[i] = 0.f; ADAPT(b)[i] = 0.f; RES(b)[i] = 1.f;
        b->spikeCount[i] = 0;
    }
    memset(b->ringCount, 0, sizeof(int32_t) * (size_t)b->nslots);
}

fbo_brain *fbo_create(int32_t N, int64_t E, const uint32_t *indptr,
                      const uint32_t *indices, const float *W,
                      const float *sign, const fbo_params *p, double seed) {
    fbo_brain *b = (fbo_brain *)calloc(1, sizeof(fbo_brain));
    if (!b) return NULL;
    b->N = N; b->E = E;
    b->nslots = fbo_nslots(p->delay, p->dt);
    fbo_constants(p, N, &b->k);
    b->indptr = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)N + 1));
    b->indices = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(E ? E : 1));
    b->W = (float *)malloc(sizeof(float) * (size_t)(E ? E : 1));
    b->sign = (float *)malloc(sizeof(float) * (size_t)N);
    memcpy(b->indptr, indptr, sizeof(uint32_t) * ((size_t)N + 1));
    if (E) {
        memcpy(b->indices, indices, sizeof(uint32_t) * (size_t)E);
        memcpy(b->W, W, sizeof(float) * (size_t)E);
    }
    memcpy(b->sign, sign, sizeof(float) * (size_t)N);
    for (int a = 0; a < FBO_NARRAYS; a++) b->arr[a] = (float *)calloc((size_t)N, sizeof(float));
    b->spikeCount = (uint32_t *)calloc((size_t)N, sizeof(uint32_t));
    b->ring = (int32_t *)calloc((size_t)b->nslots * (size_t)N, sizeof(int32_t));
    b->ringCount = (int32_t *)calloc((size_t)b->nslots, sizeof(int32_t));
    b->driven = (int32_t *)calloc((size_t)N, sizeof(int32_t));
    b->nDriven = 0;
    b->head = 0; b->lastSlot = 0; b->nFired = 0;
    b->bgAcc = 0.f;
    b->rng = fbo_seed_hash(seed);
    fbo_reset(b);
    return b;
}

void fbo_destroy(fbo_brain *b) {
    if (!b) return;
    free(b->indptr); free(b->indices); free(b->W); free(b->sign);
    for (int a = 0; a < FBO_NARRAYS; a++) free(b->arr[a]);
    free(b->spikeCount); free(b->ring); free(b->ringCount); free(b->driven);
    free(b);
}

void fbo_set_driven(fbo_brain *b, const int32_t *list, int32_t n) {
    /* lifwasm.js:70: the driven list is the drivenSet in insertion order */
    for (int32_t k = 0; k < n; k++) b->driven[k] = list[k];
    b->nDriven = n;
}

float *fbo_array(fbo_brain *b, int32_t which) {
    if (which < 0 || which >= FBO_NARRAYS) return NULL;
    return b->arr[which];
}
uint32_t *fbo_spike_count(fbo_brain *b) { return b->spikeCount; }
uint32_t fbo_rng_state(const fbo_brain *b) { return b->rng; }
void fbo_set_rng(fbo_brain *b, uint32_t rng) { b->rng = rng; }
int32_t fbo_head(const fbo_brain *b) { return b->head; }
int32_t fbo_nslots_of(const fbo_brain *b) { return b->nslots; }
float fbo_bg_acc(const fbo_brain *b) { return b->bgAcc; }

const int32_t *fbo_last_fired(const fbo_brain *b, int32_t *n) {
    *n = b->nFired;
    return b->ring + (int64_t)b->lastSlot * b->N;
}

/* Phase 1 (lif.c:34-44): spikes emitted nslots-1 steps ago reach their
 * targets, in stored list order, each row in CSR order. */
static void phase_deliver(fbo_brain *b) {
    const int32_t N = b->N;
    const int32_t *arriving = b->ring + (int64_t)b->head * N;
    const int32_t na = b->ringCount[b->head];
    float *gE = GE(b), *gI = GI(b), *res = RES(b);
    for (int32_t k = 0; k < na; k++) {
        const int32_t pre = arriving[k];
        float amp = b->sign[pre] * res[pre];
        if (amp == 0.f) continue;
        if (b->k.preInh > 0.f && DRIVE(b)[pre] > 0.f) {
            const float inh = gI[pre] < 0.f ? -gI[pre] : 0.f;
            amp = amp / (1.f + b->k.preInh * inh);
        }
        res[pre] = res[pre] - b->k.depU * res[pre];
        float *acc = amp > 0 ? gE : gI;
        for (uint32_t j = b->indptr[pre]; j < b->indptr[pre + 1]; j++) {
            const uint32_t q = b->indices[j];
            acc[q] = acc[q] + b->W[j] * amp;
        }
    }
    b->ringCount[b->head] = 0;
}

/* Record a spike: the same bookkeeping for forced and threshold spikes,
 * except the refractory value, which the caller passes. */
static inline void emit(fbo_brain *b, int32_t i, float refrValue, int32_t *out, int32_t *nf) {
    V(b)[i] = b->k.vReset;
    REFR(b)[i] = refrValue;
    out[(*nf)++] = i;
    b->spikeCount[i]++;
    TRACE(b)[i] = 1.f;
    ADAPT(b)[i] = ADAPT(b)[i] + b->k.adaptInc;
}

int32_t fbo_step(fbo_brain *b) {
    const int32_t N = b->N;
    const fbo_consts *k = &b->k;
    phase_deliver(b);

    const int32_t slot = (b->head + b->nslots - 1) % b->nslots;
    int32_t *out = b->ring + (int64_t)slot * N;
    int32_t nf = 0;

    /* Phase 1b (lif.c:46-48): background events, one draw each */
    b->bgAcc = b->bgAcc + k->bgEvents;
    const int32_t nbg = (int32_t)b->bgAcc;
    b->bgAcc = b->bgAcc - (float)nbg;
    for (int32_t e = 0; e < nbg; e++) {
        const uint32_t i = fbo_xorshift32(&b->rng) % (uint32_t)N;
        GE(b)[i] = GE(b)[i] + k->bgAmp;
    }

    /* Phase 2 (lif.c:49-52): forced spikes; a refractory neuron takes no draw */
    const float forcedRefr = k->tRef + k->dt;
    for (int32_t d = 0; d < b->nDriven; d++) {
        const int32_t i = b->driven[d];
        if (REFR(b)[i] > 0) continue;
        const float pFire = DRIVE(b)[i] * k->dtS;
        const float u = (float)(fbo_xorshift32(&b->rng) >> 8) * (1.0f / 16777216.0f);
        if (u < pFire) emit(b, i, forcedRefr, out, &nf);
    }

    /* Phase 3 (lif.c:53-72): dense membrane update and decays */
    float *v = V(b), *gE = GE(b), *gI = GI(b), *refr = REFR(b), *trace = TRACE(b),
          *adapt = ADAPT(b), *res = RES(b), *bias = BIAS(b);
    for (int32_t i = 0; i < N; i++) {
        const float vi = v[i], r = refr[i];
        float drive_mV;
        if (k->coba) {
            const float exc = (gE[i] * (k->eExc - vi)) * k->cE;
            const float inh = (gI[i] * (vi - k->eInh)) * k->cI;
            drive_mV = (((k->vRest - vi) + exc) + inh) + bias[i];
        } else {
            drive_mV = (((k->vRest - vi) + gE[i]) + gI[i]) + bias[i];
        }
        const float dv = drive_mV * k->kM;
        v[i] = r > 0 ? k->vReset : vi + dv;
        refr[i] = r > 0 ? r - k->dt : r;
        gE[i] = gE[i] * k->dE;
        gI[i] = gI[i] * k->dE;
        trace[i] = trace[i] * k->dTr;
        adapt[i] = adapt[i] * k->dA;
        res[i] = res[i] + (1.f - res[i]) * k->kRec;
    }

    /* Phase 4 (lif.c:73-76): threshold crossings in ascending index order */
    const float *thr = THR(b);
    for (int32_t i = 0; i < N; i++) {
        if (v[i] >= (k->vThresh + adapt[i]) + thr[i] && refr[i] <= 0)
            emit(b, i, k->tRef, out, &nf);
    }

    /* close the step (lif.c:77-79) */
    b->ringCount[slot] = nf;
    b->lastSlot = slot;
    b->nFired = nf;
    b->head = (b->head + 1) % b->nslots;
    return nf;
}

int64_t fbo_run(fbo_brain *b, int32_t steps, int32_t *nfired, int32_t *raster,
                int64_t cap) {
    int64_t total = 0;
    for (int32_t t = 0; t < steps; t++) {
        const int32_t nf = fbo_step(b);
        if (nfired) nfired[t] = nf;
        if (raster) {
            const int32_t *lst = b->ring + (int64_t)b->lastSlot * b->N;
            for (int32_t q = 0; q < nf; q++) {
                if (total + q < cap) raster[total + q] = lst[q];
            }
        }
        total += nf;
    }
    return total;
}

/* Apply a setter the way the reference's JS wrapper does: the new value is
 * formed in f64 and stored to the f32 array (lifwasm.js:71-76). */
void fbo_add(fbo_brain *b, int32_t which, int32_t i, double delta) {
    float *a = b->arr[which];
    a[i] = (float)((double)a[i] + delta);
}
void fbo_store(fbo_brain *b, int32_t which, int32_t i, double value) {
    b->arr[which][i] = (float)value;
}

void fbo_set_background(fbo_brain *b, double rateHz, double ampMv, double dt) {
    /* lifwasm.js:69 */
    b->k.bgEvents = (float)((double)b->N * rateHz * dt / 1000.0);
    b->k.bgAmp = (float)ampMv;
}

/* ------------------------------------------------------------------ flyvis */

/* src/wasm/lif.c:82-93: graded synapses accumulate the rectified source
 * activity into acc in source order, then a leaky update with kdt =
 * dt / max(tau, dt).  Every operation is binary32 in C order. */
void fbo_fv_step(int32_t N, const float *bias, const float *kdt, const int32_t *indptr, const int32_t *target,
                 const float *weight, float *v, float *acc, const float *x) {
    for (int32_t i = 0; i < N; i++) acc[i] = 0.f;
    for (int32_t j = 0; j < N; j++) {
        const float r = v[j];
        if (r <= 0.f) continue;
        for (int32_t k = indptr[j]; k < indptr[j + 1]; k++) {
            const float w = weight[k] * r;
            acc[target[k]] = acc[target[k]] + w;
        }
    }
    for (int32_t i = 0; i < N; i++) {
        const float d = ((-v[i] + bias[i]) + acc[i]) + x[i];
        v[i] = v[i] + kdt[i] * d;
    }
}

/* src/flyvisdiff.js:154-182: v_new = v + kdt (-v + bias + acc + x), so the
 * adjoint l of v_new gives kdt (l u), bias (l kdt), the fan-in (lacc = l kdt)
 * and v itself (l (1 - kdt)); the fan-out then passes lacc back through the
 * rectified sources, and gives each edge lacc[target] * v_source. */
void fbo_fv_back_step(int32_t N, const int32_t *indptr, const int32_t *target, const float *weight, const float *kdt,
                      const float *vPre, const float *uAt, const float *dLdV, float *lv, float *lacc, float *gw,
                      float *gb, float *gk, float *dLdLum, const int32_t *inNode, const int32_t *inCol, int32_t nIn) {
    for (int32_t i = 0; i < N; i++) {
        const double l = (double)lv[i] + (double)dLdV[i];
        if (l == 0.0) { lacc[i] = 0.f; lv[i] = 0.f; continue; }
        const double k = kdt[i];
        gk[i] = (float)((double)gk[i] + l * (double)uAt[i]);
        gb[i] = (float)((double)gb[i] + l * k);
        lacc[i] = (float)(l * k);
        lv[i] = (float)(l * (1.0 - k));
    }
    if (dLdLum)
        for (int32_t q = 0; q < nIn; q++)
            dLdLum[inCol[q]] = (float)((double)dLdLum[inCol[q]] + (double)lacc[inNode[q]]);
    for (int32_t j = 0; j < N; j++) {
        const double r = vPre[j];
        if (r <= 0.0) continue;
        double s = 0.0;
        for (int32_t k = indptr[j]; k < indptr[j + 1]; k++) {
            const double la = lacc[target[k]];
            if (la == 0.0) continue;
            s += la * (double)weight[k];
            gw[k] = (float)((double)gw[k] + la * r);
        }
        lv[j] = (float)((double)lv[j] + s);
    }
}
