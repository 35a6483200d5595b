/*
 * flybrain.h -- C ABI of the MI355X engine for fly-brain's connectome LIF step
 * and its surrogate-gradient adjoint.
 *
 * Every entry point returns an fb_status (0 = FB_OK) and never throws; the text
 * of the last failure on the calling thread is fb_last_error().  Buffers are
 * plain host pointers; device memory is owned by the library.
 *
 * Reference interfaces replaced (paths into Lulzx/fly-brain):
 *   fb_graph_*   the shared read-only connectome of src/brainsetup.js:10-19 and
 *                the per-device graph cache of src/lifgpu.js:154-181
 *   fb_brain_*,
 *   fb_step, ... LIFWasm (src/lifwasm.js:36-87) over lif_step (src/wasm/lif.c:29-80),
 *                and its WebGPU sibling LIFGpu (src/lifgpu.js:183-370)
 *   fb_batch_*   the one-brain-per-worker ensembles (src/brainsetup.js:8,
 *                scripts/lif_ensemble.mjs:16-35) and the Bend lane ensemble
 *                (kernel.bend:293-378) as one launch over B instances
 *   fb_diff_*    LIFDiff.forward / backward (src/lifdiff.js:219-686) and the
 *                WebGPU adjoint LIFAdjointGPU (src/lifdiff_gpu.js:246-503)
 *
 * Threading: one brain per host thread; a graph may be shared by any number of
 * brains and threads (it is reference counted and read-only on the device).
 */
#ifndef FLYBRAIN_H
#define FLYBRAIN_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t fb_status;
enum {
    FB_OK = 0,
    FB_ERR_ARG = 1,          /* bad argument (null, out of range, size mismatch) */
    FB_ERR_DEVICE = 2,       /* HIP runtime failure or no usable GPU */
    FB_ERR_UNSUPPORTED = 3,  /* valid request this build does not run on the GPU */
    FB_ERR_OVERFLOW = 4,     /* a bounded device buffer would overflow */
    FB_ERR_STATE = 5         /* call not valid in the object's current state */
};

/* Model parameters, as the reference's JS wrapper holds them (f64), restricted
 * to what the step reads: DEFAULTS of src/lif.js:6-15 plus the background
 * fields of src/lifwasm.js:67,69.  Derived constants (exp(-dt/tauSyn), dt/tauM,
 * 1/(eExc-vRest), ...) are formed in f64 and rounded once to f32, exactly as
 * src/lifwasm.js:58-67 writes the brain header.  coba is 0 or 1. */
typedef struct {
    double dt, vRest, vThresh, vReset, tauM, tauSyn, tRef, delay;
    double traceTau, adaptInc, adaptTau, depU, depTau;
    double coba, eExc, eInh;
    double bgRate, bgAmp, preInh;
} fb_params;

/* State arrays of one brain (the typed-array fields of LIFWasm,
 * src/lifwasm.js:43-45).  All are f32[N] except FB_SPIKECOUNT (u32[N]). */
enum {
    FB_V = 0, FB_GE, FB_GI, FB_REFR, FB_TRACE, FB_ADAPT, FB_RES,
    FB_BIAS, FB_THR, FB_DRIVE, FB_SPIKECOUNT, FB_NARRAYS
};

/* Stepping engines.  All produce identical rasters and state. */
enum {
    FB_ENGINE_AUTO = 0,        /* the fastest engine valid for the brain's parameters */
    FB_ENGINE_LAUNCH = 1,      /* one kernel sequence per step */
    FB_ENGINE_GRAPH = 2,       /* the same sequence captured in a hipGraph, replayed */
    FB_ENGINE_PERSISTENT = 3   /* one launch of a grid-barrier kernel advances many
                                  steps; the synaptic delay is used as a lookahead window */
};

typedef struct fb_graph fb_graph;
typedef struct fb_brain fb_brain;

const char *fb_last_error(void);
fb_status fb_device_info(char *name, int32_t cap, int32_t *n_cu);

/* ---------------------------------------------------------------- graph
 * CSR by presynaptic neuron: indptr u32[N+1], indices u32[E], the baked
 * weight W f32[E] and sign f32[N] that writeGraph produces (src/lifwasm.js:14-34;
 * lif.c reads them as W[j] * sign[pre] * res[pre]).  counts (i32[E], may be
 * NULL) are the raw synapse counts; they are required by the ensemble and
 * gradient engines, which rebuild W = count x gain on the device. */
fb_status fb_graph_create(int32_t N, int64_t E, const uint32_t *indptr, const uint32_t *indices,
                          const float *weights, const float *sign, const int32_t *counts,
                          fb_graph **out);
fb_status fb_graph_retain(fb_graph *g);
fb_status fb_graph_release(fb_graph *g);

/* ---------------------------------------------------------------- brain
 * new LIFWasm({graph, N, params, seed}) (src/lifwasm.js:37-54): state reset,
 * bias/thr/drive zero, RNG = (seed * 2654435761) >>> 0 || 1, ring head 0. */
fb_status fb_brain_create(fb_graph *g, const fb_params *p, double seed, fb_brain **out);
fb_status fb_brain_destroy(fb_brain *b);
/* reset(): v = vRest, gE = gI = refr = trace = adapt = 0, res = 1, spike counts
 * and ring counts zero; RNG, ring head, bias, thr, drive and the driven set are
 * kept (src/lifwasm.js:78). */
fb_status fb_reset(fb_brain *b);
/* setParams(q) (src/lif.js:47, src/lifgpu.js:349): rewrites the constants; a
 * change of delay/dt that changes nslots re-creates the delay line empty. */
fb_status fb_set_params(fb_brain *b, const fb_params *p);
fb_status fb_get_params(const fb_brain *b, fb_params *p);

/* Setters of src/lifwasm.js:69-76.  Values are f64 as JS numbers are; a
 * setter stores f32(value) and an adder stores f32(f64(old) + value). */
fb_status fb_set_drive(fb_brain *b, const int32_t *idx, int32_t n, double hz);
fb_status fb_set_drive_one(fb_brain *b, int32_t i, double hz);
fb_status fb_set_bias(fb_brain *b, const int32_t *idx, int32_t n, double mv);
fb_status fb_set_thr(fb_brain *b, int32_t i, double mv);
fb_status fb_add_g(fb_brain *b, int32_t i, double e, double ii);
fb_status fb_pulse(fb_brain *b, const int32_t *idx, int32_t n, double mv);
fb_status fb_set_background(fb_brain *b, double hz, double mv);
/* A queue of setter calls applied in order, with one upload and one launch:
 * element q acts as fb_set_drive_one (FB_DELTA_DRIVE), fb_set_bias with one
 * index (FB_DELTA_BIAS), fb_set_thr (FB_DELTA_THR), or fb_add_g adding to one
 * conductance (FB_DELTA_GE, FB_DELTA_GI) would on idx[q] with val[q].  The
 * batched form of LIFGpu's delta queue (src/lifgpu.js:140-148, 321-346):
 * callers that set thousands of thresholds per frame (src/sim/neuromod.js
 * modulate()) pay one launch instead of one per call. */
enum { FB_DELTA_DRIVE = 0, FB_DELTA_BIAS = 1, FB_DELTA_THR = 2, FB_DELTA_GE = 3, FB_DELTA_GI = 4 };
fb_status fb_apply_deltas(fb_brain *b, int32_t n, const int32_t *idx, const int32_t *kind, const double *val);
/* fb_apply_deltas with no allocation, device or host.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops).  At most 65,536 deltas a call (LIFGpu's
 * queue bound); more is an FB_ERR_ARG.  The staging is one fixed array per
 * device, held until the process ends and shared by every call on that
 * device: calls that overlap on one GPU race. */
fb_status fb_apply_deltas_noalloc(fb_brain *b, int32_t n, const int32_t *idx, const int32_t *kind, const double *val);
/* the driven set in insertion order (drivenSet of src/lifwasm.js:51,70-71) */
fb_status fb_driven(const fb_brain *b, int32_t *out, int32_t cap, int32_t *n);

/* step() x nsteps (src/lifwasm.js:79-85 over lif_step).  nfired_last may be NULL. */
fb_status fb_step(fb_brain *b, int32_t nsteps, int32_t engine, int32_t *nfired_last);
/* The spikes of the last step, in lif.c's order: forced (driven) spikes in
 * driven-set order, then threshold spikes in ascending index (lif.c:49-76). */
fb_status fb_fired(fb_brain *b, int32_t *out, int32_t cap, int32_t *n);
/* Runs nsteps and records every step's spike count and spike list.
 * raster receives the lists back to back (at most cap entries); total is the
 * number of spikes in the run even if it exceeds cap. */
fb_status fb_run_record(fb_brain *b, int32_t nsteps, int32_t engine, int32_t *nfired,
                        int32_t *raster, int64_t cap, int64_t *total);
fb_status fb_get_array(fb_brain *b, int32_t which, void *dst);
fb_status fb_set_array(fb_brain *b, int32_t which, const void *src);
/* RNG state, ring head, background accumulator, simulated time in ms */
fb_status fb_brain_scalars(fb_brain *b, uint32_t *rng, int32_t *head, float *bg_acc, double *t_ms);
/* The brain's neuron count: the length of every fb_get_array / fb_set_array
 * buffer.  No counterpart in the reference (its typed arrays carry their
 * length); for bindings that check a caller's buffer before a copy. */
fb_status fb_brain_size(const fb_brain *b, int32_t *N);
/* The xorshift32 state itself (non-zero): a lane of the cord ensemble starts
 * from a raw state, 2654435761 + 7919 m + c (cordens.bend:296-300), not from
 * LIFWasm's seed hash. */
fb_status fb_set_rng(fb_brain *b, uint32_t rng);
/* The closed proprioceptive loop of the cord ensemble (cordens.bend:227-273):
 * after every step the step's spikes are tallied by pool (poolOf[N]: pool
 * 0..nPools-1, any larger value none; nPools <= 8), the counts recorded (up
 * to recordCap steps), and the drive of targets[] for the next step set to
 * base + gain * (the pools' total `delay` steps back), in binary32, the total
 * of a step before the first being 0.  The drive values change, the driven
 * list does not.  Runs on the per-step engines; poolOf NULL turns it off. */
fb_status fb_set_pool_feedback(fb_brain *b, const uint8_t *poolOf, int32_t nPools, const int32_t *targets,
                               int32_t nTargets, float base, float gain, int32_t delay, int32_t recordCap);
/* the recorded per-step pool counts, [n][nPools] uint16, n = *nRecords (all
 * steps since the loop was set; at most cap copied) */
fb_status fb_pool_counts(fb_brain *b, uint16_t *dst, int32_t cap, int32_t *nRecords);
/* Work done since create: edge events (the sum of out-degrees of the arriving
 * spikes, the reference's "edges/ms" of scripts/profile_fly.mjs:19,25),
 * synapse events (the same sum weighted by synapse count; needs counts at
 * graph creation), spikes emitted and steps run. */
fb_status fb_brain_counters(fb_brain *b, int64_t *edge_events, int64_t *synapse_events, int64_t *spikes,
                            int64_t *steps);
fb_status fb_synchronize(fb_brain *b);
/* Runs the brain's work on the caller's HIP stream (a hipStream_t passed as a
 * pointer; NULL restores the brain's own stream), so a caller can order it with
 * its own work and time it with its own events.  No counterpart in the
 * reference, which has no streams. */
fb_status fb_brain_set_stream(fb_brain *b, void *stream);
/* Engine statistics by name: "fallback_steps" (steps the lookahead engine
 * handed to the per-step engine after a buffer overflow), "persistent_launches",
 * "persistent_kernel_ms" (summed HIP-event time of those launches),
 * "persistent_supported" (1 if the brain's parameters allow the lookahead
 * engine), "nslots", "serial_deliver".  No counterpart in the reference. */
fb_status fb_brain_stat(fb_brain *b, const char *name, double *value);

/* ---------------------------------------------------------------- ensembles
 * B <= 64 members of one brain model over one shared graph, advanced together
 * (the one-brain-per-worker ensembles of src/brainsetup.js:8 and
 * scripts/lif_ensemble.mjs:16-35; the lane ensemble of kernel.bend:293-378).
 * Member b is exactly the brain fb_brain_create would make with seed seeds[b]
 * and weights W[j] = f32(count[j] * gains[b*nTypes + typeOf[pre(j)]]) (the
 * graph needs counts); all members share the driven set and drive rates.
 * Members stop, state frozen and no draws, after their own step count
 * (fb_batch_set_steps; unset = run forever). */
typedef struct fb_batch fb_batch;
fb_status fb_batch_create(fb_graph *g, const fb_params *p, int32_t B, const double *seeds, const int32_t *typeOf,
                          int32_t nTypes, const float *gains, fb_batch **out);
fb_status fb_batch_destroy(fb_batch *b);
fb_status fb_batch_set_drive(fb_batch *b, const int32_t *idx, int32_t n, double hz);
/* per-member FB_THR or FB_BIAS on a list of neurons; silencing is thr = 1e6
 * (setThr(i, 1e6) of scripts/lif_ensemble.mjs:35) */
fb_status fb_batch_set_member(fb_batch *b, int32_t which, int32_t member, const int32_t *idx, int32_t n, double value);
/* steps[b] more steps for member b (negative: no limit) */
fb_status fb_batch_set_steps(fb_batch *b, const int32_t *steps);
/* Advances nsteps.  With nactive != NULL records, per step, the neurons that
 * spiked in any member (nactive[s] of them) with their spike mask (bit b =
 * member b) and forced mask (bit b = the spike was a forced one), appended to
 * ids/masks/fmasks up to cap entries; total = all entries of the run. */
fb_status fb_batch_run(fb_batch *b, int32_t nsteps, int32_t *nactive, int32_t *ids, uint64_t *masks,
                       uint64_t *fmasks, int64_t cap, int64_t *total);
fb_status fb_batch_get_array(fb_batch *b, int32_t which, int32_t member, void *dst);
/* Lane mode, for the lane ensemble of cordens.bend:140-160: signGain = 1
 * delivers W[j] * f32(sign * gain[b][type]) (the member's type gain scales
 * the presynaptic sign; the graph's baked W is used) instead of
 * f32(count * gain) * sign; laneDrive [B][N] gives every member its own
 * drive rates over the shared driven list (NULL: the shared rates). */
fb_status fb_batch_set_lanes(fb_batch *b, int32_t signGain, const float *laneDrive);
/* each member's xorshift32 state itself, states[B] (non-zero) */
fb_status fb_batch_set_rng(fb_batch *b, const uint32_t *states);
/* fb_set_pool_feedback for every member at once, with its own delay[B]; the
 * loop writes the per-lane drive (fb_batch_set_lanes first) */
fb_status fb_batch_set_pool_feedback(fb_batch *b, const uint8_t *poolOf, int32_t nPools, const int32_t *targets,
                                     int32_t nTargets, float base, float gain, const int32_t *delay, int32_t recordCap);
/* dst [B][n][nPools] uint16, n = min(cap, *nRecords) */
fb_status fb_batch_pool_counts(fb_batch *b, uint16_t *dst, int32_t cap, int32_t *nRecords);
fb_status fb_batch_rng(fb_batch *b, uint32_t *rng);
/* neurons, members and types: the lengths the per-member arrays, gains and
 * per-member lists above take.  No counterpart in the reference; for bindings
 * that check a caller's buffer before a copy. */
fb_status fb_batch_size(const fb_batch *b, int32_t *N, int32_t *B, int32_t *nTypes);
fb_status fb_batch_counters(fb_batch *b, int64_t *edge_events, int64_t *synapse_events, int64_t *spikes,
                            int64_t *steps);
fb_status fb_batch_set_stream(fb_batch *b, void *stream);
/* FB_ENGINE_AUTO (default: the lookahead engine when the ensemble allows it
 * and a run covers at least two delay windows), FB_ENGINE_LAUNCH (per-step
 * kernels) or FB_ENGINE_PERSISTENT (the lookahead engine, or
 * FB_ERR_UNSUPPORTED).  Both give identical results.  No counterpart in the
 * reference. */
fb_status fb_batch_set_engine(fb_batch *b, int32_t engine);
/* The lookahead engine's regular-block shape, the tile the ensemble is cut
 * into: 72, 80, 96 or 128 neurons (x all members) per workgroup, or 0 (default)
 * for the autotuner, which runs the first epochs of the first run long
 * enough on every shape that fits the device, twice round, and keeps the
 * fastest; later ensembles of the same graph, size and driven set take that
 * choice without measuring (0 measures again).  Every shape gives the same
 * bits.  "lookahead_shape" and
 * "autotune_us_per_epoch_<72|80|96|128>" (fb_batch_stat) report the choice and
 * its measurements.  No counterpart in the reference. */
fb_status fb_batch_set_shape(fb_batch *b, int32_t neuronsPerBlock);
/* "lookahead_launches", "lookahead_kernel_ms", "lookahead_steps",
 * "fallback_steps" (steps handed to the per-step kernels after a bucket
 * overflow), "lookahead_supported" */
fb_status fb_batch_stat(fb_batch *b, const char *name, double *value);

/* ---------------------------------------------------------------- differentiable model
 * The reference's LIFDiff (src/lifdiff.js:84-686) on the GPU: a COBA network
 * with per-neuron log-gain, type gain (outScale), threshold offset, size
 * scaling (inScale from sizeLog), masked threshold term (kcThreshold) and bias
 * (laminaBias), depression and adaptation; a hard threshold with the
 * fast-sigmoid surrogate 1/(beta (1 + |u - thr|/beta)^2) in the backward pass,
 * or a logistic threshold ("soft") that makes the forward pass smooth.
 * Arithmetic follows LIFDiff: doubles for every intermediate, binary32 for
 * every stored array.  fb_diff_forward records a tape (the spike lists) and a
 * checkpoint every `checkpoint` steps; fb_diff_backward replays one segment at
 * a time and sweeps it in reverse (src/lifdiff.js:358-464, 494-686).
 *
 * Parameters are LIFDiff's options in this field order (lifdiff.js:88-112);
 * `checkpoint` is the segment length K (the reference fixes 16). */
typedef struct {
    double dt, vRest, vThresh, vReset, tauM, tauSyn, tRef, delay;
    double wSyn, adaptInc, adaptTau, depU, depTau;
    double eExc, eInh, inhGain, kcThreshold, laminaBias;
    double surrogateBeta, adjClip;
    double sizeAlpha, maxSizeScale, boostCap;
    int32_t soft, checkpoint;
    /* exact32 (lifdiff.js:105-109, 173-190, 231-240, 263-274, 317-326): the
     * forward pass in lif.c's binary32 arithmetic on writeGraph's baked
     * weights and signs, bit-identical to lif_step.  logGain and the spike
     * amplitude must be inert and thrMask empty (forward refuses otherwise);
     * the depression of res rounds as lif.c does.  The backward pass and its
     * replay are LIFDiff's binary64 model either way, as in the reference. */
    int32_t exact32;
} fb_diff_params;

/* per-neuron inputs (fb_diff_set_input).  FB_D_INSCALE is a caller-given
 * inScale taken verbatim instead of the one formed from sizeLog (opts.inScale,
 * lifdiff.js:115-118, 160); FB_D_BETA a per-neuron surrogateBeta
 * (lifdiff.js:89-95, 335, 590).  A NULL array clears either of the two. */
enum { FB_D_OUTSCALE = 0, FB_D_LOGGAIN, FB_D_THROFFSET, FB_D_SIZELOG, FB_D_DRIVE, FB_D_INSCALE_IN, FB_D_BETA,
       FB_D_NINPUTS };
/* state arrays (fb_diff_get_state) */
enum { FB_D_V = 0, FB_D_GE, FB_D_GI, FB_D_ADAPT, FB_D_RES, FB_D_REFR, FB_D_SPIKECOUNT, FB_D_INSCALE, FB_D_NSTATE };
/* the nine scalar gradients, in DIFF_PARAMS order (src/lifdiff.js:75) */
enum { FB_G_WSYN = 0, FB_G_SIZEALPHA, FB_G_INHGAIN, FB_G_EINH, FB_G_VTHRESH, FB_G_KC, FB_G_LAMINA,
       FB_G_ADAPTINC, FB_G_DEPU, FB_G_NPARAMS };

typedef struct fb_diff fb_diff;
/* new LIFDiff(data, opts) (lifdiff.js:84): CSR rows in the caller's edge
 * order (gradients per edge come back in that order), `counts` the delivered
 * synapse count per edge with the minSyn / sensory gate folded in as zero
 * (lifdiff.js:167-172), `sign` +1 / -1 / 0 per neuron (EXC_SIGN of its
 * transmitter, or preSign). */
fb_status fb_diff_create(int32_t N, int64_t E, const uint32_t *indptr, const uint32_t *indices, const float *counts,
                         const float *sign, const fb_diff_params *p, fb_diff **out);
fb_status fb_diff_destroy(fb_diff *d);
fb_status fb_diff_set_params(fb_diff *d, const fb_diff_params *p);
fb_status fb_diff_get_params(fb_diff *d, fb_diff_params *p);
/* outScale, logGain, thrOffset, sizeLog, drive (Hz), inScale or
 * surrogateBeta: N values */
fb_status fb_diff_set_input(fb_diff *d, int32_t which, const float *values);
/* per-edge multiplier on the delivered weight (opts.edgeGain, lifdiff.js:126-129); NULL = all 1 */
fb_status fb_diff_set_edge_gain(fb_diff *d, const float *edgeGain);
fb_status fb_diff_set_masks(fb_diff *d, const uint8_t *thrMask, const uint8_t *biasMask);
/* the driven set in insertion order (setDrive / setDriveOne, lifdiff.js:196-207) */
fb_status fb_diff_set_driven(fb_diff *d, const int32_t *idx, int32_t n);
/* dLdSpike[N] and lossFrom of backward() (lifdiff.js:494); NULL = zero */
fb_status fb_diff_set_loss(fb_diff *d, const float *dLdSpike, int32_t lossFrom);
/* A filtered-rate loss on R readouts, added to the flat term: r_t = a r_{t-1}
 * + (1-a) s_t / (dt/1000), a = exp(-dt/tau), L = mean over steps and readouts
 * of (r_t - target)^2.  No counterpart in the reference. */
fb_status fb_diff_set_readout(fb_diff *d, const int32_t *idx, int32_t R, const float *targetHz, double tauMs);
/* tape entries allowed per step on average (default max(4096, N/4)) */
fb_status fb_diff_set_tape_capacity(fb_diff *d, int64_t perStep);
/* driveIdx / driveEpoch (src/lifdiff.js:135-144, 228-234): the drive of the
 * neurons idx[0..nD) is an input.  At the start of epoch e (every epochSteps
 * steps) their rates become sched[e*nD + k], through setDriveOne's semantics:
 * a rate of 0 takes the neuron out of the driven set, a positive rate puts it
 * back at the end (the onEpoch hook of LIFDiff.forward); -0.0 stands for a
 * positive rate too small for binary32, which joins the set with a zero drive.  The schedule must
 * cover the forward pass.  driveSoft (src/lifdiff.js:338): a driven neuron
 * spikes with the drive's expectation sd + (1 - sd) st instead of a Poisson
 * draw.  nD = 0 clears. */
fb_status fb_diff_set_drive_schedule(fb_diff *d, const int32_t *idx, int32_t nD, int32_t epochSteps,
                                     const float *sched, int32_t nEpochs, int32_t driveSoft);
/* dLdSpikePerEpoch (src/lifdiff.js:519-522): the loss weight of a spike of
 * neuron i in epoch e is dPer[e*N + i] (replacing dLdSpike); NULL clears.
 * Like dLdSpike it may be set between the forward and the backward pass (the
 * loss the forward pass reported is then the old one's). */
fb_status fb_diff_set_loss_per_epoch(fb_diff *d, const float *dPer, int32_t nEpochs, int32_t epochSteps);
/* dL/d sched after fb_diff_backward: [nEpochs][nD] (the `drive` result of
 * LIFDiff.backward, src/lifdiff.js:490-492) */
fb_status fb_diff_drive_grad(fb_diff *d, float *gDrive);
/* forward(steps, {spikeTrain}) (src/lifdiff.js:219, 243, 298-307; the replay
 * at 432-438): at step t < steps the neurons idx[ptr[t] .. ptr[t+1]) are
 * forced to spike, in that order, after the step's drive draws and without a
 * draw -- a neuron refractory by then is skipped -- and join the step's list
 * after its driven spikes.  ptr has steps + 1 entries from 0; steps = 0
 * clears.  It applies to the forward passes that follow; a recorded tape
 * keeps the forced spikes it ran with. */
fb_status fb_diff_set_spike_train(fb_diff *d, int32_t steps, const int64_t *ptr, const int32_t *idx);
/* forward(steps, {seed, record}) (lifdiff.js:219): from reset; loss = the
 * flat term sum dLdSpike * spike count from lossFrom on, plus the readout term */
fb_status fb_diff_forward(fb_diff *d, int32_t steps, double seed, int32_t record, double *loss);
/* backward(tape, dLdSpike, {truncate, lossFrom}) (lifdiff.js:494): params[9]
 * in FB_G_* order, gLogGain[N], gEdge[E] = dL/d edgeGain (no counterpart in
 * the reference), gInScale[N]; clamped[3] = non-finite values zeroed among
 * params, logGain and edges.  Any output may be NULL. */
fb_status fb_diff_backward(fb_diff *d, int32_t truncate, double *params, float *gLogGain, float *gEdge,
                           float *gInScale, int32_t *clamped);
/* the results of the last backward pass (fb_diff_backward or a batch's),
 * outputs as fb_diff_backward's; any may be NULL */
fb_status fb_diff_grads(fb_diff *d, double *params, float *gLogGain, float *gEdge, float *gInScale,
                        int32_t *clamped);
/* dL/d gain_T for outScale[i] = gain of i's type: sum over the type of
 * gLogGain[i] / outScale[i] (the type gains of src/brainmodel.js:63-100) */
fb_status fb_diff_type_grads(fb_diff *d, const int32_t *typeOf, int32_t nTypes, double *gType);
fb_status fb_diff_get_state(fb_diff *d, int32_t which, float *dst);
/* the tape: off[steps+1] and entries (forced spikes have bit 31 of idx set) */
fb_status fb_diff_spikes(fb_diff *d, int64_t *off, int32_t *idx, float *amp, int64_t cap, int64_t *total);
/* readout spike amplitudes s[steps][R] and rates[steps][R] of the last forward */
fb_status fb_diff_readout(fb_diff *d, float *s, double *rate);
fb_status fb_diff_rng(fb_diff *d, uint32_t *rng);
/* neurons, edges, steps of the recorded forward pass and readouts: the
 * lengths the inputs, gradients, tape and readout arrays take.  No
 * counterpart in the reference; for bindings that check a caller's buffer. */
fb_status fb_diff_size(const fb_diff *d, int32_t *N, int64_t *E, int32_t *steps, int32_t *R);
fb_status fb_diff_set_stream(fb_diff *d, void *stream);
/* "forward_ms", "backward_ms", "edge_events", "arrivals", "steps", "tape_entries" */
fb_status fb_diff_stat(fb_diff *d, const char *name, double *value);

/* Several fb_diff networks (same delay length and checkpoint interval, each
 * with its own inputs) stepped in lockstep: every step's kernels run once for
 * all members (the member is the launch's second grid dimension), member b
 * for its own steps[b].  The batched forward + adjoint of many instances over
 * one connectome (one network per worker in the reference,
 * src/brainsetup.js:8, scripts/grad_fit.mjs).  No counterpart in the
 * reference's API. */
typedef struct fb_diff_batch fb_diff_batch;
fb_status fb_diff_batch_create(fb_diff *const *members, int32_t B, fb_diff_batch **out);
fb_status fb_diff_batch_destroy(fb_diff_batch *b);
/* forward passes: member b runs steps[b] steps from reset with seeds[b] */
fb_status fb_diff_batch_forward(fb_diff_batch *b, const int32_t *steps, const double *seeds, int32_t record,
                                double *losses);
/* backward passes of the last batch forward; read each member's gradients
 * with fb_diff_grads / fb_diff_type_grads */
fb_status fb_diff_batch_backward(fb_diff_batch *b, int32_t truncate);
/* "forward_ms", "backward_ms" of the last batch passes */
fb_status fb_diff_batch_stat(fb_diff_batch *b, const char *name, double *value);

/* ------------------------------------------------------------------ flyvis
 * The optic-lobe rate network that drives the connectome from the eyes
 * (src/flyvis.js:3-25 over fv_step, src/wasm/lif.c:82-93): passive point
 * neurons with graded synapses,
 *   v_i += kdt_i (-v_i + bias_i + sum_{j: v_j > 0} w_ji v_j + x_i),
 *   kdt = dt / max(tau, dt),
 * bit-identical to fv_step.  A model (graph by source: indptr[N+1],
 * target[E], weight[E]; bias, tau per node) is shared by its instances, one
 * per eye (the `shared` parts of src/flyvis.js:12). */
typedef struct fb_fv_model fb_fv_model;
typedef struct fb_fv fb_fv;
enum { FB_FV_V = 0, FB_FV_ACC = 1 };
fb_status fb_fv_model_create(int32_t N, int64_t E, const int32_t *indptr, const int32_t *target, const float *weight,
                             const float *bias, const float *tau, double dt, fb_fv_model **out);
fb_status fb_fv_model_release(fb_fv_model *m);
/* a new instance: v = bias, no input (new FlyVis, src/flyvis.js:4-19) */
fb_status fb_fv_create(fb_fv_model *m, fb_fv **out);
fb_status fb_fv_destroy(fb_fv *e);
/* x[N]: the input of every node (setInput writes luminance into R1-R8) */
fb_status fb_fv_set_input(fb_fv *e, const float *x);
fb_status fb_fv_step(fb_fv *e, int32_t nsteps);
fb_status fb_fv_get(fb_fv *e, int32_t which, float *dst);
fb_status fb_fv_set_v(fb_fv *e, const float *v);
/* v = bias (reset(), src/flyvis.js:24) */
fb_status fb_fv_reset(fb_fv *e);

/* The differentiable flyvis (FlyVisDiff, src/flyvisdiff.js:36-185): the same
 * step, a tape and the reverse sweep.
 * kdt given directly instead of from tau (model.kdt, src/flyvisdiff.js:50-53) */
fb_status fb_fv_model_set_kdt(fb_fv_model *m, const float *kdt);
/* the photoreceptor map, flattened type by type as src/flyvisdiff.js:60-64
 * does: input node inNode[q] reads hex column inCol[q]; nCol = max + 1 */
fb_status fb_fv_model_set_input_map(fb_fv_model *m, int32_t nIn, const int32_t *inNode, const int32_t *inCol);
/* lum[nCol]: x[inNode[q]] = lum[inCol[q]] in map order (setInput,
 * src/flyvisdiff.js:75-81) */
fb_status fb_fv_set_lum(fb_fv *e, const float *lum);
/* startTape(steps) (src/flyvisdiff.js:104-107): room for `steps` taped steps,
 * none taped yet */
fb_status fb_fv_tape_start(fb_fv *e, int32_t steps);
/* stepTaped() nsteps times (src/flyvisdiff.js:110-114); refuses to tape past
 * the room fb_fv_tape_start made */
fb_status fb_fv_step_taped(fb_fv *e, int32_t nsteps);
fb_status fb_fv_tape_len(fb_fv *e, int32_t *n);
/* Nodes, edges and luminance columns of an instance's model: the lengths
 * fb_fv_set_input / set_v / get (N), fb_fv_backward (E, tape x N, tape x
 * nCol) and fb_fv_set_lum (nCol) copy.  No counterpart in the reference; for
 * bindings that check a caller's buffer before a copy. */
fb_status fb_fv_size(const fb_fv *e, int32_t *N, int64_t *E, int32_t *nCol);
/* Nodes and edges of a model (the length fb_fv_model_set_kdt reads is N). */
fb_status fb_fv_model_size(const fb_fv_model *m, int32_t *N, int64_t *E);
/* backward(dLdV, g, dLdLum) (src/flyvisdiff.js:128-185) over the n taped
 * steps: dLdV[n][N] the loss gradient of v after each step; gW[E] (source
 * order), gBias[N], gKdt[N] accumulators, added to (two eyes share one);
 * dLdLum[n][nCol] accumulator or NULL.  Host pointers. */
fb_status fb_fv_backward(fb_fv *e, const float *dLdV, float *gW, float *gBias, float *gKdt, float *dLdLum);

#ifdef __cplusplus
}
#endif
#endif
