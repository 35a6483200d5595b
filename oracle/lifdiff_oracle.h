/*
 * lifdiff_oracle.h -- declarations for the CPU restatement of the reference's
 * differentiable LIF (lifdiff_oracle.c).  TEST INFRASTRUCTURE ONLY: tests/ and
 * the bench's CPU baseline use it as the checker; the product never links it.
 */
#ifndef FLYBRAIN_LIFDIFF_ORACLE_H
#define FLYBRAIN_LIFDIFF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* LIFDiff model options (src/lifdiff.js:88-112), f64 as JS holds them, in the
 * order of the product's fb_diff_params struct (include/flybrain.h). */
typedef struct {
    double dt, vRest, vThresh, vReset, tauM, tauSyn, tRef, delay;
    double wSyn, adaptInc, adaptTau, depU, depTau;
    double eExc, eInh, inhGain, kcThreshold, laminaBias;
    double surrogateBeta, adjClip;
    double sizeAlpha, maxSizeScale, boostCap;
    int32_t soft, checkpoint;
    int32_t exact32;   /* lif.c's binary32 arithmetic in the forward pass (src/lifdiff.js:105-109) */
} fdo_params;

/* per-neuron inputs.  FDO_INSCALE_IN is a caller-given inScale taken verbatim
 * instead of the one formed from sizeLog (opts.inScale, src/lifdiff.js:115-118,
 * 160); FDO_BETA a per-neuron surrogateBeta (src/lifdiff.js:89-95).  A NULL
 * value array clears either of these two. */
enum {
    FDO_OUTSCALE = 0, FDO_LOGGAIN, FDO_THROFFSET, FDO_SIZELOG, FDO_DRIVE, FDO_INSCALE_IN, FDO_BETA, FDO_NINPUTS
};
/* state arrays after forward() */
enum {
    FDO_V = 0, FDO_GE, FDO_GI, FDO_ADAPT, FDO_RES, FDO_REFR, FDO_SPIKECOUNT, FDO_INSCALE, FDO_NSTATE
};
/* order of the nine scalar gradients (src/lifdiff.js:75) */
enum {
    FDO_G_WSYN = 0, FDO_G_SIZEALPHA, FDO_G_INHGAIN, FDO_G_EINH, FDO_G_VTHRESH,
    FDO_G_KC, FDO_G_LAMINA, FDO_G_ADAPTINC, FDO_G_DEPU, FDO_NPARAMS
};

typedef struct fdo_net fdo_net;

/* counts: synapse counts per edge (the minSyn / sensory gate already folded in
 * as a zero); sign: +1 / -1 / 0 per presynaptic neuron (EXC_SIGN of its
 * transmitter). */
fdo_net *fdo_create(int32_t N, int64_t E, const uint32_t *indptr, const uint32_t *indices,
                    const float *counts, const float *sign, const fdo_params *p);
void fdo_destroy(fdo_net *n);
void fdo_set_input(fdo_net *n, int32_t which, const float *values);
void fdo_set_edge_gain(fdo_net *n, const float *edgeGain);
void fdo_set_masks(fdo_net *n, const uint8_t *thrMask, const uint8_t *biasMask);
void fdo_set_driven(fdo_net *n, const int32_t *list, int32_t count);
void fdo_set_loss(fdo_net *n, const float *dLdSpike, int32_t lossFrom);
void fdo_set_readout(fdo_net *n, const int32_t *idx, int32_t R, const float *targetHz, double tauMs);
/* driveIdx / driveEpoch: at the start of epoch e (every epochSteps steps) the
 * rate of idx[k] becomes sched[e*nD + k] through setDriveOne's semantics (the
 * last epoch's rates hold beyond nEp).  A rate of -0.0 stands for a positive
 * rate that binary32 holds as zero (it joins the driven set with a zero
 * drive).  driveSoft: the drive spikes as its expectation.  NULL idx clears
 * the schedule. */
void fdo_set_drive_schedule(fdo_net *n, const int32_t *idx, int32_t nD, int32_t epochSteps, const float *sched,
                            int32_t nEp, int32_t driveSoft);
/* dLdSpikePerEpoch: the loss weight of neuron i's spikes in epoch e is
 * dPer[e*N + i] (NULL clears) */
void fdo_set_loss_per_epoch(fdo_net *n, const float *dPer, int32_t nEp, int32_t epochSteps);
/* dL/d sched [nEp][nD] of the last backward */
void fdo_drive_grad(const fdo_net *n, float *gDrive);
void fdo_set_spike_train(fdo_net *n, int32_t steps, const int64_t *ptr, const int32_t *idx);
/* Runs `steps` steps from reset with the given seed; records the tape when
 * `record` is nonzero.  Returns the loss (NaN when exact32 is set with a
 * non-empty thrMask, which LIFDiff refuses). */
double fdo_forward(fdo_net *n, int32_t steps, double seed, int32_t record);
/* Reverse sweep over the recorded tape.  params: 9 doubles in FDO_G_* order;
 * gLogGain[N], gEdge[E] (dL/d edgeGain), gInScale[N]; clamped[3]. */
void fdo_backward(fdo_net *n, int32_t truncate, double *params, float *gLogGain, float *gEdge,
                  float *gInScale, int32_t *clamped);
float *fdo_state(fdo_net *n, int32_t which);
uint32_t fdo_rng(const fdo_net *n);
/* the recorded spike lists: off[steps+1], entries idx/amp (forced spikes have
 * bit 31 of idx set); returns the total count */
int64_t fdo_spikes(const fdo_net *n, int64_t *off, int32_t *idx, float *amp, int64_t cap);
/* readout spike amplitudes [steps][R] and filtered rates [steps][R] */
int64_t fdo_readout(const fdo_net *n, float *s, double *rate);

#ifdef __cplusplus
}
#endif
#endif
