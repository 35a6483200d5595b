// fb_diff.hip -- the differentiable model on the GPU: forward with a tape and
// checkpoints, segment replay into an arena, and the reverse sweep.
//
// The model is the reference's LIFDiff (src/lifdiff.js:84-686): COBA membrane,
// per-neuron log-gain, type gain (outScale), threshold offset, size scaling
// (inScale), a masked threshold term and bias, depression, adaptation, a hard
// or logistic ("soft") threshold, and the fast-sigmoid surrogate in the
// backward pass.  LIFDiff evaluates in doubles and stores into Float32Arrays;
// these kernels do the same (f64 locals, one rounding per stored value, in the
// reference's evaluation order), and deliver each step's arrivals with the
// ordered-sum scheme of the forward engine, so a step's conductance receives
// its arrivals in list order exactly as the serial loop adds them.
//
// Per forward step (t):
//   kd_deliver   wave per arrival (the spikes of step t - D from the tape):
//                res read and depressed, the row's addends (double) go into
//                per-(sign class, target) accumulators keyed by arrival rank
//                (driven position for a forced spike, N + index otherwise)
//                and position in row.
//   kd_driven    one workgroup: the drive draws in set order from the serial
//                xorshift32 stream (leapfrog-parallel, as in the forward
//                engine), forced spikes appended to the tape in set order.
//   kd_membrane  dense pass: ordered sums, integrate, decay, threshold, spike
//                amplitude, tape append (forward) or arena record (replay).
// Per backward step (t, reverse):
//   kd_bneuron   dense pass: reversed decays, mode dispatch, surrogate, the
//                adjoint state update; scalar terms reduced per workgroup.
//   kd_barrival  wave per arrival: the fan-out gather of lgE / lgI into lRaw,
//                per-edge gradient, gInScale scatter, gLogGain, the spike
//                adjoint of the emitting step, depression adjoint.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "fb_internal.h"

namespace fb {

uint32_t seed_hash(double seed);
fb_status leap_tables(uint32_t **jump, uint32_t **leap);

constexpr uint32_t FORCED_BIT = 0x80000000u;
constexpr int DIFF_BLOCK = 256;

struct DiffColl {
    uint64_t key;
    double addend;
    int32_t next;
    int32_t pad;
};

struct DiffDev {
    int32_t N, S, D, K, R, nDriven, soft, lossFrom;
    int64_t E;
    double dt, vRest, vThresh, vReset, wSyn, adaptInc, depU, eExc, eInh, inhGain, kc, lamina, beta, aClip;
    double dE, dA, kRec, kM, dtS, cE, cI, emitThr, sizeAlpha, boostCap, minScale;
    float forcedRefr, tRefF;
    // graph and inputs
    const uint32_t *indptr, *indices;
    const float *counts, *edgeGain;   // edgeGain null = every edge at 1
    const float *sign, *outScale, *thrOffset, *inScale, *drive, *sizeLog, *dLdSpike;
    const double *expLogGain;
    const uint8_t *thrMask, *biasMask;
    const int32_t *driven, *drvPos, *readIdx;
    // state
    float *v, *gE, *gI, *adapt, *res, *refr, *spikeCount, *scored;
    uint32_t *rng;
    uint8_t *forced;
    int32_t *dflag;
    // ordered delivery
    int32_t *cnt, *chain;
    double *add;
    uint64_t *key;
    DiffColl *coll;
    int64_t collHalf;
    int32_t *ncoll;                   // [2], by step parity
    int32_t *longList, *nLong;        // accumulators past DIFF_LONG arrivals this step, and their count
    // tape
    int64_t *off;                     // [steps + 1]
    unsigned long long *top;
    int32_t *idx;
    float *amp;
    int64_t cap;
    float *readS;                     // [steps][R]
    int32_t *err;
    unsigned long long *events;       // [0] edge events, [1] arrivals
    const uint32_t *leap;
    // arena, K steps
    float *aV, *aGE, *aGI, *aAd, *aSp, *resAt;
    uint8_t *aMode;
    // adjoint
    float *lv, *lgE, *lgI, *lad, *lres, *ring, *gLogGain, *gEdge, *gInScale;
    double *accW, *accDepU, *accInh, *blockAcc;
    double *rate, *dLds;               // [steps][R]
    const float *target;
    float *ckpt;                       // checkpoints [n][6N + 1]
    // the drive as an input (driveIdx / driveEpoch): rates per epoch of `ep`
    // steps for the neurons with drvSlot >= 0, and the drive gradient
    int32_t nDI, ep, nEp, driveSoft;
    const int32_t *drvSlot;            // [N] position in driveIdx, or -1; null without a schedule
    const float *drvSched;             // [nEp][nDI]
    float *gDrive;                     // [nEp][nDI]
    const float *dPer;                 // [epochs][N] per-epoch spike weights, or null
    double *lossAcc;                   // [N] the per-epoch weighted loss of a forward pass
    const float *betaArr;              // [N] per-neuron surrogateBeta, or null (the scalar)
    // the spike train (forward's spikeTrain): step t < trainSteps forces the
    // neurons trainIdx[trainPtr[t] .. trainPtr[t + 1]) in that order
    int32_t trainSteps;
    const int64_t *trainPtr;
    const int32_t *trainIdx;
    // exact32: the forward pass in lif.c's binary32 arithmetic -- every
    // scalar as lif.c reads it from its binary32 header, and writeGraph's
    // weights and signs baked from the inputs before each forward pass
    int32_t exact32;
    float xvRest, xeExc, xeInh, xvThresh, xvReset, xtRef, xadaptInc, xdt, xdepU, xlamina;
    float xdE, xdA, xkRec, xkM, xdtS, xcE, xcI, xforcedRefr;
    float *W32, *signW32;              // [E], [N]
};

constexpr double DRIVE_MAX = 0.999;   // src/lifdiff.js:72

// The drive rate of neuron i at step t: the schedule's rate for its epoch
// (held past the last one) if i is in driveIdx, else the fixed input.
__device__ __forceinline__ float drive_at(const DiffDev &d, int32_t i, int32_t t) {
    if (d.drvSlot) {
        const int32_t k = d.drvSlot[i];
        if (k >= 0) {
            int32_t e = t / d.ep;
            if (e >= d.nEp) e = d.nEp - 1;
            return d.drvSched[(int64_t)e * d.nDI + k];
        }
    }
    return d.drive[i];
}

// the loss weight of a spike of neuron i at step t (per epoch when dPer is set)
__device__ __forceinline__ double spike_weight(const DiffDev &d, int32_t i, int32_t t) {
    return d.dPer ? (double)d.dPer[(int64_t)(t / d.ep) * d.N + i] : (double)d.dLdSpike[i];
}

enum { DERR_COLL = 1, DERR_TAPE = 2, DERR_DRAWS = 4 };

// ------------------------------------------------------------------ helpers

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    return x;
}

// g + the arrivals of one accumulator, in key order, each rounded to f32 as a
// Float32Array += does
__device__ inline float ordered_sum_d(float g, int32_t count, uint64_t key0, double add0, int32_t head,
                                      const DiffColl *coll) {
    if (count <= 32) {
        uint64_t ks[32];
        double as[32];
        int n = 0;
        ks[n] = key0; as[n] = add0; n++;
        for (int32_t p = head; p >= 0 && n < 32; p = coll[p].next) {
            const uint64_t kk = coll[p].key;
            const double aa = coll[p].addend;
            int m = n++;
            while (m > 0 && ks[m - 1] > kk) { ks[m] = ks[m - 1]; as[m] = as[m - 1]; m--; }
            ks[m] = kk; as[m] = aa;
        }
        for (int m = 0; m < n; m++) g = (float)((double)g + as[m]);
        return g;
    }
    uint64_t last = 0;
    bool first = true;
    for (int32_t c = 0; c < count; c++) {
        uint64_t best = ~0ull;
        double ba = 0.0;
        if ((first || key0 > last) && key0 < best) { best = key0; ba = add0; }
        for (int32_t p = head; p >= 0; p = coll[p].next) {
            const uint64_t kk = coll[p].key;
            if ((first || kk > last) && kk < best) { best = kk; ba = coll[p].addend; }
        }
        g = (float)((double)g + ba);
        last = best;
        first = false;
    }
    return g;
}

// An accumulator with more arrivals than DIFF_LONG in one step is summed by
// dx_long before the membrane pass: its arrivals are staged in LDS (a walk of
// its chain), sorted by key with a bitonic network and added in that order --
// linear in the arrivals where the selection above is quadratic.
constexpr int DIFF_LONG = 32, DIFF_LONG_CAP = 2048, DIFF_LONG_WG = 128;

__device__ __forceinline__ void dx_long(const DiffDev &d) {
    __shared__ uint64_t ks[DIFF_LONG_CAP];
    __shared__ double as[DIFF_LONG_CAP];
    __shared__ int32_t s_m;
    const int32_t N = d.N, nL = *d.nLong, tid = threadIdx.x;
    for (int32_t p = blockIdx.x; p < nL; p += gridDim.x) {
        const int32_t ci = d.longList[p];
        const int32_t n = d.cnt[ci];
        float *gp = ci < N ? d.gE + ci : d.gI + (ci - N);
        if (n > DIFF_LONG_CAP) {                 // beyond the stage: the selection, one thread
            if (tid == 0) *gp = ordered_sum_d(*gp, n, d.key[ci], d.add[ci], d.chain[ci], d.coll);
        } else {
            if (tid == 0) {
                int32_t m = 0;
                ks[m] = d.key[ci];
                as[m] = d.add[ci];
                m++;
                for (int32_t q = d.chain[ci]; q >= 0 && m < n; q = d.coll[q].next) {
                    ks[m] = d.coll[q].key;
                    as[m] = d.coll[q].addend;
                    m++;
                }
                s_m = m;                          // fewer than n only after a flagged pool overflow
            }
            __syncthreads();
            const int32_t m = s_m;
            int32_t P = 1;
            while (P < m) P <<= 1;
            for (int32_t q = m + tid; q < P; q += blockDim.x) ks[q] = ~0ull;
            __syncthreads();
            for (int32_t k = 2; k <= P; k <<= 1)
                for (int32_t j = k >> 1; j > 0; j >>= 1) {
                    for (int32_t q = tid; q < P; q += blockDim.x) {
                        const int32_t o = q ^ j;
                        if (o > q) {
                            const bool up = (q & k) == 0;
                            if ((ks[q] > ks[o]) == up) {
                                const uint64_t tk = ks[q]; ks[q] = ks[o]; ks[o] = tk;
                                const double ta = as[q]; as[q] = as[o]; as[o] = ta;
                            }
                        }
                    }
                    __syncthreads();
                }
            if (tid == 0) {
                float g = *gp;
                for (int32_t q = 0; q < m; q++) g = (float)((double)g + as[q]);
                *gp = g;
            }
        }
        if (tid == 0) {
            d.cnt[ci] = 0;
            d.chain[ci] = -1;
        }
        __syncthreads();
    }
}

__device__ __forceinline__ float clip_to(double x, double c) {
    if (c > 0) x = x > c ? c : (x < -c ? -c : x);
    return (float)x;
}

// ------------------------------------------------------------------ forward / replay

// The arrivals of step t: the tape's spikes of step t - D.  resBase is the
// tape index of the replay segment's first arrival.
template <bool REPLAY>
__device__ __forceinline__ void dx_deliver(const DiffDev &d, int32_t t, int32_t resFromT, int32_t nblk) {
    if (t < d.D) return;
    const int lane = threadIdx.x & 63;
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)nblk * blockDim.x) >> 6;
    // with a one-step delay the list being read is the previous step's, whose
    // end the tape counter still holds (its offset is written by kd_driven)
    // (an overflowed tape, already flagged, is read only up to its capacity)
    const int64_t a0 = min(d.off[t - d.D], d.cap);
    const int64_t a1 = min((!REPLAY && d.D == 1) ? (int64_t)*d.top : d.off[t - d.D + 1], d.cap);
    const int64_t resBase = REPLAY ? d.off[resFromT] : 0;
    if (!REPLAY && d.exact32) {
        // lif.c's delivery (src/lifdiff.js:263-274): s = f32(sign * res), a zero
        // s skipped before the depression, which rounds as lif.c's
        // res -= depU * res; one binary32 product per edge (zero ones skipped:
        // adding them changes no conductance)
        for (int64_t k = a0 + wave; k < a1; k += nwaves) {
            const uint32_t e = (uint32_t)d.idx[k];
            const int32_t pre = (int32_t)(e & ~FORCED_BIT);
            const float resv = d.res[pre];
            float s = d.signW32[pre] * resv;
            const float ampk = d.amp[k];
            if (ampk != 1.f) s = s * ampk;
            if (s == 0.f) continue;
            const uint32_t ra = d.indptr[pre], re = d.indptr[pre + 1];
            if (lane == 0) {
                d.res[pre] = resv - d.xdepU * resv;
                atomicAdd(&d.events[0], (unsigned long long)(re - ra));
                atomicAdd(&d.events[1], 1ull);
            }
            const uint32_t rank = (e & FORCED_BIT) ? (uint32_t)(k - a0) : (uint32_t)d.N + (uint32_t)pre;
            const int32_t cls = s > 0 ? 0 : d.N;
            const int32_t par = t & 1;
            for (uint32_t j = ra + lane; j < re; j += 64) {
                const float dlt = d.W32[j] * s;
                if (dlt == 0.f) continue;
                const uint32_t q = d.indices[j];
                const int32_t ci = cls + (int32_t)q;
                const uint64_t kk = ((uint64_t)rank << 32) | (uint64_t)(j - ra);
                const int32_t slot = atomicAdd(&d.cnt[ci], 1);
                if (slot == DIFF_LONG) d.longList[atomicAdd(d.nLong, 1)] = ci;
                if (slot == 0) {
                    d.add[ci] = (double)dlt;
                    d.key[ci] = kk;
                } else {
                    const int32_t p = atomicAdd(&d.ncoll[par], 1);
                    if (p < d.collHalf) {
                        const int32_t gp = (int32_t)(par * d.collHalf + p);
                        d.coll[gp].key = kk;
                        d.coll[gp].addend = (double)dlt;
                        d.coll[gp].next = atomicExch(&d.chain[ci], gp);
                    } else {
                        atomicOr(d.err, DERR_COLL);
                    }
                }
            }
        }
        return;
    }
    for (int64_t k = a0 + wave; k < a1; k += nwaves) {
        const uint32_t e = (uint32_t)d.idx[k];
        const int32_t pre = (int32_t)(e & ~FORCED_BIT);
        const float sg = d.sign[pre];
        if (sg == 0.f) continue;
        const float resv = d.res[pre];
        const double eff = (double)sg * d.wSyn * (double)d.outScale[pre] * d.expLogGain[pre] * (double)resv *
                           (double)d.amp[k];
        const uint32_t ra = d.indptr[pre], re = d.indptr[pre + 1];
        if (lane == 0) {
            if (REPLAY) d.resAt[k - resBase] = resv;
            d.res[pre] = (float)((double)resv - d.depU * (double)resv);
            if (!REPLAY) {
                atomicAdd(&d.events[0], (unsigned long long)(re - ra));
                atomicAdd(&d.events[1], 1ull);
            }
        }
        // arrival order: forced spikes first, in the driven-set order of their
        // step (the pre-pass wrote them at the head of the step's list), then
        // threshold spikes by index (the membrane pass appends them in any order)
        const uint32_t rank = (e & FORCED_BIT) ? (uint32_t)(k - a0) : (uint32_t)d.N + (uint32_t)pre;
        const bool exc = eff > 0;
        const int32_t cls = exc ? 0 : d.N;
        const int32_t par = t & 1;
        for (uint32_t j = ra + lane; j < re; j += 64) {
            const uint32_t q = d.indices[j];
            const double egj = d.edgeGain ? (double)d.edgeGain[j] : 1.0;
            const double addend = exc ? (double)d.counts[j] * egj * (double)d.inScale[q] * eff
                                      : (double)d.counts[j] * egj * (double)d.inScale[q] * d.inhGain * eff;
            const int32_t ci = cls + (int32_t)q;
            const uint64_t kk = ((uint64_t)rank << 32) | (uint64_t)(j - ra);
            const int32_t slot = atomicAdd(&d.cnt[ci], 1);
            if (slot == DIFF_LONG) d.longList[atomicAdd(d.nLong, 1)] = ci;   // once per accumulator a step
            if (slot == 0) {
                d.add[ci] = addend;
                d.key[ci] = kk;
            } else {
                const int32_t p = atomicAdd(&d.ncoll[par], 1);
                if (p < d.collHalf) {
                    const int32_t gp = (int32_t)(par * d.collHalf + p);
                    d.coll[gp].key = kk;
                    d.coll[gp].addend = addend;
                    d.coll[gp].next = atomicExch(&d.chain[ci], gp);
                } else {
                    atomicOr(d.err, DERR_COLL);
                }
            }
        }
    }
}

// A driven neuron's draw u against its rate (src/lifdiff.js:293): binary64
// u < drive dt/1000, or under exact32 lif.c's binary32 threshold
__device__ __forceinline__ bool drive_fires_at(const DiffDev &d, float rate, double u) {
    return d.exact32 ? u < (double)(rate * d.xdtS) : u < (double)rate * d.dtS;
}
__device__ __forceinline__ bool drive_fires(const DiffDev &d, int32_t i, int32_t t, double u) {
    return drive_fires_at(d, drive_at(d, i, t), u);
}

// a forced spike's refractory period and adaptation increment (src/lifdiff.js:294-295)
__device__ __forceinline__ void forced_effects(const DiffDev &d, int32_t i) {
    if (d.exact32) {
        d.refr[i] = d.xforcedRefr;
        d.adapt[i] = d.adapt[i] + d.xadaptInc;
    } else {
        d.refr[i] = d.forcedRefr;
        d.adapt[i] = (float)((double)d.adapt[i] + d.adaptInc);
    }
}

// The spike train's step t (src/lifdiff.js:298-307), after the drive
// pre-pass and by the same workgroup: each listed neuron, in list order, is
// forced like a driven spike without a draw unless it is refractory by then
// (drawn this step, listed before, or still refractory), and goes to the tape
// after the step's driven spikes.  The train is a few neurons a step: thread
// 0 walks it.
__device__ __forceinline__ void dx_train(const DiffDev &d, int32_t t) {
    if (t >= d.trainSteps) return;
    __syncthreads();   // the pre-pass's refractory writes and the tape top
    if (threadIdx.x != 0) return;
    unsigned long long top = *d.top;
    for (int64_t k = d.trainPtr[t]; k < d.trainPtr[t + 1]; k++) {
        const int32_t i = d.trainIdx[k];
        if (d.refr[i] > 0) continue;
        forced_effects(d, i);
        d.forced[i] = 1;
        d.spikeCount[i] = (float)((double)d.spikeCount[i] + 1.0);
        if (t >= d.lossFrom) d.scored[i] = (float)((double)d.scored[i] + 1.0);
        if (d.dPer && t >= d.lossFrom) d.lossAcc[i] += spike_weight(d, i, t);
        const int32_t rk = d.readIdx[i];
        if (rk >= 0) d.readS[(int64_t)t * d.R + rk] = 1.f;
        if ((int64_t)top < d.cap) {
            d.idx[top] = (int32_t)((uint32_t)i | FORCED_BIT);
            d.amp[top] = 1.f;
        } else {
            atomicOr(d.err, DERR_TAPE);
        }
        top++;
    }
    *d.top = top;
}

// The drive pre-pass (src/lifdiff.js:290-297): one draw per non-refractory
// driven neuron in set order; forced spikes go to the tape in that order.
__device__ __forceinline__ void dx_driven(const DiffDev &d, int32_t t) {
    constexpr int T = 1024;
    __shared__ uint32_t leapS[LEAP_LDS_WORDS];
    __shared__ uint32_t ybase[64];
    __shared__ uint32_t draws[LEAP_ROUND];
    __shared__ int32_t scan[64 + T / 64];
    __shared__ uint32_t s_rng;
    __shared__ unsigned long long s_top;
    const int tid = threadIdx.x;
    if (tid == 0) {
        s_rng = *d.rng;
        s_top = *d.top;
        d.off[t] = (int64_t)s_top;      // this step's list starts here
    }
    leap_load<T>(leapS, d.leap);
    __syncthreads();
    uint32_t base = s_rng;
    const int32_t nD = d.nDriven;
    const int32_t per = (nD + T - 1) / T;
    const int32_t lo = min(nD, tid * per), hi = min(nD, lo + per);
    int32_t mine = 0;
    for (int32_t q = lo; q < hi; q++)
        if (!(d.refr[d.driven[q]] > 0)) mine++;
    int32_t totalDraws;
    const int32_t off = wg_exclusive_scan<T>(mine, scan, &totalDraws);
    for (int32_t q2 = lo; q2 < hi; q2++) d.dflag[q2] = 0;
    int32_t q = lo, next = off, fired = 0;
    for (int32_t r0 = 0; r0 < totalDraws; r0 += LEAP_ROUND) {
        const int32_t cnt = min(LEAP_ROUND, totalDraws - r0);
        leap_round(leapS, ybase, draws, cnt, base);
        while (next < off + mine && next < r0 + cnt) {
            const int32_t i = d.driven[q];
            if (!(d.refr[i] > 0)) {
                const double u = (double)(draws[next - r0] >> 8) * (1.0 / 16777216.0);
                if (drive_fires(d, i, t, u)) {
                    forced_effects(d, i);
                    d.forced[i] = 1;
                    d.spikeCount[i] = (float)((double)d.spikeCount[i] + 1.0);
                    if (t >= d.lossFrom) d.scored[i] = (float)((double)d.scored[i] + 1.0);
                    if (d.dPer && t >= d.lossFrom) d.lossAcc[i] += spike_weight(d, i, t);
                    const int32_t rk = d.readIdx[i];
                    if (rk >= 0) d.readS[(int64_t)t * d.R + rk] = 1.f;
                    d.dflag[q] = 1;
                    fired++;
                }
                next++;
            }
            q++;
        }
        base = draws[cnt - 1];
        __syncthreads();
    }
    {
        int32_t totalFired;
        int32_t pos = wg_exclusive_scan<T>(fired, scan, &totalFired);
        for (int32_t q2 = lo; q2 < hi; q2++)
            if (d.dflag[q2]) {
                const int64_t k = (int64_t)s_top + pos++;
                if (k < d.cap) {
                    d.idx[k] = (int32_t)((uint32_t)d.driven[q2] | FORCED_BIT);
                    d.amp[k] = 1.f;
                } else {
                    atomicOr(d.err, DERR_TAPE);
                }
            }
        if (tid == 0) *d.top = s_top + (unsigned long long)totalFired;
    }
    if (tid == 0) *d.rng = base;
    dx_train(d, t);
}

// The drive pre-pass with at most 2048 driven neurons (one workgroup of 1024
// threads): each thread holds its at most two consecutive driven neurons in
// registers, finds its first stream position with a scan of non-refractory
// flags, and jumps there in two leapfrog products, M^(32 (off/32)) then
// M^(off mod 32 + 1); the draws after it are single steps.  The fine tables
// (M^1..M^32) go to LDS and the 64 coarse products M^(32 h) x of the step's
// stream state are formed while the driven neurons' refractory state and
// rates load, so that after the scan a jump is one LDS product.  (Loading the
// state a forced spike updates for every driven neuron up front costs more
// than it saves: those are scattered loads from one workgroup.)  Same draws,
// effects and tape order as dx_driven.
__device__ __forceinline__ void dx_driven_fast(const DiffDev &d, int32_t t) {
    constexpr int T = 1024, DP = 2;
    __shared__ int32_t scan[2][T / 64];
    __shared__ unsigned long long s_top;
    __shared__ uint32_t s_rng;
    __shared__ uint32_t fineS[32 * LEAP_STRIDE];
    __shared__ uint32_t ybase[64];
    const int tid = threadIdx.x;
    leap_load<T, 32>(fineS, d.leap);
    if (tid == 0) {
        s_rng = *d.rng;
        s_top = *d.top;
        d.off[t] = (int64_t)s_top;      // this step's list starts here
    }
    const int32_t nD = d.nDriven, P = (nD + T - 1) / T;
    int32_t ii[DP];
    float rate[DP];
    uint32_t nr = 0;   // bit u: driven neuron u of this thread is not refractory and draws
#pragma unroll
    for (int u = 0; u < DP; u++) {
        const int32_t q = tid * P + u;
        ii[u] = (u < P && q < nD) ? d.driven[q] : -1;
        rate[u] = ii[u] >= 0 ? drive_at(d, ii[u], t) : 0.f;
        if (ii[u] >= 0 && !(d.refr[ii[u]] > 0)) nr |= 1u << u;
    }
    __syncthreads();   // s_rng, s_top and the fine tables
    if (tid < 64) ybase[tid] = matvec(d.leap + (32 + tid) * LEAP_WORDS, s_rng);
    int32_t n;
    const int32_t off = wg_scan_flags<T, DP>(nr, scan[0], &n);   // its barrier publishes ybase
    uint32_t fm = 0;   // bit u: neuron u was forced to spike
    if (nr) {
        uint32_t x = matvec(fineS + (off & 31) * LEAP_STRIDE, ybase[off >> 5]);
        int32_t j = off;
        bool first = true;
#pragma unroll
        for (int u = 0; u < DP; u++) {
            if (!((nr >> u) & 1u)) continue;
            if (!first) x = xs32(x);
            first = false;
            const int32_t i = ii[u];
            const double uu = (double)(x >> 8) * (1.0 / 16777216.0);
            if (drive_fires_at(d, rate[u], uu)) {
                forced_effects(d, i);
                d.forced[i] = 1;
                d.spikeCount[i] = (float)((double)d.spikeCount[i] + 1.0);
                if (t >= d.lossFrom) d.scored[i] = (float)((double)d.scored[i] + 1.0);
                if (d.dPer && t >= d.lossFrom) d.lossAcc[i] += spike_weight(d, i, t);
                const int32_t rk = d.readIdx[i];
                if (rk >= 0) d.readS[(int64_t)t * d.R + rk] = 1.f;
                fm |= 1u << u;
            }
            if (j == n - 1) *d.rng = x;   // the step's last draw is the stream state after it
            j++;
        }
    }
    int32_t nf;
    int32_t pos = wg_scan_flags<T, DP>(fm, scan[1], &nf);
#pragma unroll
    for (int u = 0; u < DP; u++) {
        if (!((fm >> u) & 1u)) continue;
        const int64_t k = (int64_t)s_top + pos++;
        if (k < d.cap) {
            d.idx[k] = (int32_t)((uint32_t)ii[u] | FORCED_BIT);
            d.amp[k] = 1.f;
        } else {
            atomicOr(d.err, DERR_TAPE);
        }
    }
    if (tid == 0) *d.top = s_top + (unsigned long long)nf;
    dx_train(d, t);
}

// The drive pre-pass when replaying a segment: the forward put every forced
// spike of step t on the tape (FORCED_BIT, at the head of the step's list),
// so the replay applies their effects without drawing the stream again.
__device__ __forceinline__ void dx_forced_replay(const DiffDev &d, int32_t t) {
    const int64_t a0 = min(d.off[t], d.cap), a1 = min(d.off[t + 1], d.cap);
    for (int64_t k = a0 + threadIdx.x; k < a1; k += blockDim.x) {
        const uint32_t e = (uint32_t)d.idx[k];
        if (!(e & FORCED_BIT)) continue;
        const int32_t i = (int32_t)(e & ~FORCED_BIT);
        d.refr[i] = d.forcedRefr;
        d.adapt[i] = (float)((double)d.adapt[i] + d.adaptInc);
        d.forced[i] = 1;
    }
}

// Integrate, decay, threshold (src/lifdiff.js:313-345; replay 422-457).
template <bool REPLAY>
__device__ __forceinline__ void dx_membrane(const DiffDev &d, int32_t t, int32_t r) {
    const int32_t N = d.N;
    const int32_t i = blockIdx.x * DIFF_BLOCK + threadIdx.x;
    bool emit = false;
    float ampOut = 0.f;
    if (i < N) {
        float gEf = d.gE[i], gIf = d.gI[i];
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const int32_t ci = c * N + i;
            const int32_t n = d.cnt[ci];
            if (n) {
                float g = c ? gIf : gEf;
                if (n == 1) {
                    g = (float)((double)g + d.add[ci]);
                } else {
                    g = ordered_sum_d(g, n, d.key[ci], d.add[ci], d.chain[ci], d.coll);
                    d.chain[ci] = -1;
                }
                d.cnt[ci] = 0;
                if (c) gIf = g; else gEf = g;
            }
        }
        const float vf = d.v[i];
        double vi = vf, s = 0.0;
        float refrf = d.refr[i], adaptf = d.adapt[i], resf = d.res[i];
        const float refr0 = refrf, adapt0 = adaptf, res0 = resf;
        const bool wasRefr = refrf > 0;
        const bool x32 = !REPLAY && d.exact32;
        uint8_t md = 0;
        const int64_t ar = (int64_t)r * N + i;
        if (REPLAY) { d.aV[ar] = vf; d.aGE[ar] = gEf; d.aGI[ar] = gIf; }
        if (wasRefr) {
            if (x32) {
                refrf = refrf - d.xdt;
                vi = d.xvReset;
            } else {
                refrf = (float)((double)refrf - d.dt);
                vi = d.vReset;
            }
            if (REPLAY) {
                if (d.forced[i]) { md = 2; s = 1.0; } else md = 1;
            }
        } else if (x32) {
            // lif.c's COBA expression, one binary32 rounding per operation in
            // its order (src/lifdiff.js:317-326)
            const float bias = d.biasMask[i] ? d.xlamina : 0.f;
            float acc = d.xvRest - vf;
            acc = acc + (gEf * (d.xeExc - vf)) * d.xcE;
            acc = acc + (gIf * (vf - d.xeInh)) * d.xcI;
            vi = (double)(vf + (acc + bias) * d.xkM);
        } else {
            const double bias = d.biasMask[i] ? d.lamina : 0.0;
            vi += (d.vRest - vi + (double)gEf * (d.eExc - vi) * d.cE + (double)gIf * (vi - d.eInh) * d.cI + bias) * d.kM;
        }
        if (x32) {
            gEf = gEf * d.xdE;
            gIf = gIf * d.xdE;
            adaptf = adaptf * d.xdA;
            resf = resf + (1.f - resf) * d.xkRec;
        } else {
            gEf = (float)((double)gEf * d.dE);
            gIf = (float)((double)gIf * d.dE);
            adaptf = (float)((double)adaptf * d.dA);
            resf = (float)((double)resf + (1.0 - (double)resf) * d.kRec);
        }
        if (REPLAY) d.aAd[ar] = adaptf;
        if (!wasRefr) {
            const double thr = x32 ? (double)((d.xvThresh + adaptf) + d.thrOffset[i])
                                   : d.vThresh + (double)adaptf + (d.thrMask[i] ? d.kc : 0.0) + (double)d.thrOffset[i];
            const double beta = d.betaArr ? (double)d.betaArr[i] : d.beta;
            s = d.soft ? 1.0 / (1.0 + exp(-(vi - thr) / beta)) : (vi >= thr ? 1.0 : 0.0);
            // the drive as an expectation (src/lifdiff.js:338, 453)
            if (d.driveSoft) {
                const float dv = drive_at(d, i, t);
                if (dv > 0) {
                    const double sd = fmin(DRIVE_MAX, (double)dv * d.dtS);
                    s = sd + (1.0 - sd) * s;
                    md = 3;
                }
            }
            if (s > 0) {
                vi = s * (x32 ? (double)d.xvReset : d.vReset) + (1.0 - s) * vi;
                if (!d.soft) refrf = x32 ? d.xtRef : d.tRefF;
                adaptf = (float)((double)adaptf + s * (x32 ? (double)d.xadaptInc : d.adaptInc));
            }
            if (!REPLAY) {
                // adding s = 0 changes neither count (they start at +0 and
                // only grow), so a neuron without a spike reads and writes
                // neither
                if (s != 0.0) {
                    d.spikeCount[i] = (float)((double)d.spikeCount[i] + s);
                    if (t >= d.lossFrom) d.scored[i] = (float)((double)d.scored[i] + s);
                }
                if (d.dPer && t >= d.lossFrom) d.lossAcc[i] += spike_weight(d, i, t) * s;
                const int32_t rk = d.readIdx[i];
                if (rk >= 0) d.readS[(int64_t)t * d.R + rk] = (float)s;
                emit = s > d.emitThr;
                ampOut = (float)s;
            }
        }
        if (REPLAY) { d.aMode[ar] = md; d.aSp[ar] = (float)s; }
        // a forced neuron is refractory here (its pre-pass set refr > 0); the
        // slowly changing arrays are written only where a value changed
        if (wasRefr) d.forced[i] = 0;
        d.v[i] = (float)vi;
        d.gE[i] = gEf; d.gI[i] = gIf;
        if (__float_as_uint(adaptf) != __float_as_uint(adapt0)) d.adapt[i] = adaptf;
        if (__float_as_uint(resf) != __float_as_uint(res0)) d.res[i] = resf;
        if (__float_as_uint(refrf) != __float_as_uint(refr0)) d.refr[i] = refrf;
    }
    if (!REPLAY) {
        const uint64_t m = __ballot(emit);
        const int lane = threadIdx.x & 63;
        unsigned long long base = 0;
        if (lane == 0 && m) base = atomicAdd(d.top, (unsigned long long)__popcll(m));
        base = __shfl(base, 0, 64);
        if (emit) {
            const int64_t k = (int64_t)base + __popcll(m & ((1ull << lane) - 1ull));
            if (k < d.cap) {
                d.idx[k] = i;
                d.amp[k] = ampOut;
            } else {
                atomicOr(d.err, DERR_TAPE);
            }
        }
    }
    // the other parity's collision pool was consumed by the previous step;
    // dx_long has consumed this step's long list
    if (i == 0) {
        d.ncoll[(t + 1) & 1] = 0;
        *d.nLong = 0;
    }
}

// exact32's writeGraph bake (src/lifdiff.js:182-189): wave per row, W[k] =
// f32(count * edgeGain * inScale[q] * ig) with ig = (sign < 0 ? inhGain : 1) *
// outScale[j] and sign[j] = f32(sign * wSyn), the products in binary64
__global__ void kd_bake_x32(DiffDev d) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t j = wave; j < d.N; j += nwaves) {
        const double sg = d.sign[j];
        if (lane == 0) d.signW32[j] = (float)(sg * d.wSyn);
        const double ig = (sg < 0 ? d.inhGain : 1.0) * (double)d.outScale[j];
        for (uint32_t k = d.indptr[j] + lane; k < d.indptr[j + 1]; k += 64)
            d.W32[k] = (float)((double)d.counts[k] * (d.edgeGain ? (double)d.edgeGain[k] : 1.0) *
                               (double)d.inScale[d.indices[k]] * ig);
    }
}

__global__ void kd_reset(DiffDev d, float vRestF) {
    const int32_t N = d.N;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        d.v[i] = vRestF; d.gE[i] = 0.f; d.gI[i] = 0.f; d.adapt[i] = 0.f; d.res[i] = 1.f; d.refr[i] = 0.f;
        d.spikeCount[i] = 0.f; d.scored[i] = 0.f; d.forced[i] = 0;
        d.cnt[i] = 0; d.cnt[i + N] = 0; d.chain[i] = -1; d.chain[i + N] = -1;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        d.ncoll[0] = 0; d.ncoll[1] = 0; *d.top = 0ull; d.off[0] = 0; *d.err = 0; *d.nLong = 0;
    }
}

__global__ void kd_finish(DiffDev d, int32_t steps) {
    if (threadIdx.x == 0 && blockIdx.x == 0) d.off[steps] = (int64_t)*d.top;
}

// Restores a checkpoint's state (replay start) and clears the step buffers.
__device__ __forceinline__ void dx_restore(const DiffDev &d, int32_t seg) {
    const int32_t N = d.N;
    const float *ck = d.ckpt + (int64_t)seg * (6 * (int64_t)N + 1);
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        d.v[i] = ck[i]; d.gE[i] = ck[N + i]; d.gI[i] = ck[2 * N + i]; d.adapt[i] = ck[3 * N + i];
        d.res[i] = ck[4 * N + i]; d.refr[i] = ck[5 * N + i];
        d.forced[i] = 0;
        d.cnt[i] = 0; d.cnt[i + N] = 0; d.chain[i] = -1; d.chain[i + N] = -1;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        d.ncoll[0] = 0; d.ncoll[1] = 0; *d.nLong = 0;
        *d.rng = __float_as_uint(ck[6 * N]);
    }
}

__device__ __forceinline__ void dx_save(const DiffDev &d, int32_t seg) {
    const int32_t N = d.N;
    float *ck = d.ckpt + (int64_t)seg * (6 * (int64_t)N + 1);
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        ck[i] = d.v[i]; ck[N + i] = d.gE[i]; ck[2 * N + i] = d.gI[i]; ck[3 * N + i] = d.adapt[i];
        ck[4 * N + i] = d.res[i]; ck[5 * N + i] = d.refr[i];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) ck[6 * N] = __uint_as_float(*d.rng);
}

// ------------------------------------------------------------------ losses

// sum dLdSpike * scored counts, deterministic (one workgroup)
__global__ void __launch_bounds__(1024) kd_loss_flat(DiffDev d, double *out) {
    __shared__ double red[16];
    double acc = 0.0;
    for (int32_t i = threadIdx.x; i < d.N; i += 1024)
        acc += d.dPer ? d.lossAcc[i] : (double)d.dLdSpike[i] * (double)d.scored[i];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < 16; w++) s += red[w];
        out[0] = s;
    }
}

// Filtered rates of the readouts and the loss's spike adjoint: one thread per
// readout, forward filter then reverse filter.  Per-readout loss sums go to
// part[k]; the caller adds them in order.
__global__ void kd_readout(DiffDev d, int32_t T, double a, double *part) {
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d.R) return;
    const int32_t R = d.R;
    const double tgt = (double)d.target[k];
    double r = 0.0, L = 0.0;
    for (int32_t t = 0; t < T; t++) {
        r = a * r + (1.0 - a) * (double)d.readS[(int64_t)t * R + k] / d.dtS;
        d.rate[(int64_t)t * R + k] = r;
        const double e = r - tgt;
        L += e * e;
    }
    part[k] = L;
    double lam = 0.0;
    const double TR = (double)T * (double)R;
    for (int32_t t = T - 1; t >= 0; t--) {
        lam = 2.0 * (d.rate[(int64_t)t * R + k] - tgt) / TR + a * lam;
        d.dLds[(int64_t)t * R + k] = (1.0 - a) / d.dtS * lam;
    }
}

// ------------------------------------------------------------------ backward

__device__ __forceinline__ void dx_bclear(const DiffDev &d, int32_t ringToo) {
    const int32_t N = d.N;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        d.lv[i] = 0.f; d.lgE[i] = 0.f; d.lgI[i] = 0.f; d.lad[i] = 0.f; d.lres[i] = 0.f;
        if (ringToo)
            for (int32_t s = 0; s < d.S; s++) d.ring[(int64_t)s * N + i] = 0.f;
    }
}

__global__ void kd_bzero(DiffDev d, int32_t nblk) {
    const int32_t N = d.N;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < d.E; q += stride) d.gEdge[q] = 0.f;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += stride) {
        d.gLogGain[i] = 0.f; d.gInScale[i] = 0.f; d.accW[i] = 0.0; d.accDepU[i] = 0.0; d.accInh[i] = 0.0;
    }
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < 5 * (int64_t)nblk; q += stride)
        d.blockAcc[q] = 0.0;
}

// Reversed decays and integrate-and-spike of step t (src/lifdiff.js:555-617).
__device__ __forceinline__ void dx_bneuron(const DiffDev &d, int32_t t, int32_t r) {
    __shared__ double red[5][DIFF_BLOCK / 64];
    const int32_t N = d.N;
    const int32_t i = blockIdx.x * DIFF_BLOCK + threadIdx.x;
    double pA = 0, pV = 0, pK = 0, pL = 0, pEI = 0;
    if (i < N) {
        const int64_t ar = (int64_t)r * N + i;
        const uint8_t md = d.aMode[ar];
        const double lpA = d.lad[i];
        const double lE = (double)d.lgE[i] * d.dE, lI = (double)d.lgI[i] * d.dE, lA = lpA * d.dA;
        const double lR = (double)d.lres[i] * (1.0 - d.kRec);
        float lgEf = (float)lE, lgIf = (float)lI, ladf = (float)lA;
        const float lresf = (float)lR;
        float lvf = d.lv[i];
        float *ls = d.ring + (int64_t)(t % d.S) * N + i;
        double dl = t >= d.lossFrom ? spike_weight(d, i, t) : 0.0;
        const int32_t rk = d.readIdx[i];
        if (rk >= 0) dl += d.dLds[(int64_t)t * d.R + rk];
        // this neuron's entry of the drive gradient for the epoch of t
        const int32_t slot = d.drvSlot ? d.drvSlot[i] : -1;
        float *gDr = slot >= 0 ? d.gDrive + (int64_t)min(t / d.ep, d.nEp - 1) * d.nDI + slot : nullptr;
        if (md == 1) {
            lvf = 0.f;
        } else if (md == 2) {
            // the drive fired: the adaptation increment and the drive's own
            // gradient remain (src/lifdiff.js:565-574)
            pA += lA;
            if (gDr) *gDr = (float)((double)*gDr + (lA * d.adaptInc + (double)*ls + dl) * d.dtS);
            lvf = 0.f;
            *ls = 0.f;
        } else {
            const double vv = d.aV[ar], gEv = d.aGE[ar], gIv = d.aGI[ar];
            const double bias = d.biasMask[i] ? d.lamina : 0.0;
            const double u = vv + (d.vRest - vv + gEv * (d.eExc - vv) * d.cE + gIv * (vv - d.eInh) * d.cI + bias) * d.kM;
            const double thr = d.vThresh + (double)d.aAd[ar] + (d.thrMask[i] ? d.kc : 0.0) + (double)d.thrOffset[i];
            const double dd = u - thr;
            const double s = d.aSp[ar];
            // with the soft drive the amplitude is sd + (1 - sd) st, and the
            // surrogate is taken at st (src/lifdiff.js:580-585)
            double sd = 0.0, st = s;
            if (md == 3) {
                sd = fmin(DRIVE_MAX, (double)drive_at(d, i, t) * d.dtS);
                st = (s - sd) / (1.0 - sd);
            }
            const double beta = d.betaArr ? (double)d.betaArr[i] : d.beta;
            double sg;
            if (d.soft) {
                sg = st * (1.0 - st) / beta;
            } else {
                const double q = 1.0 + fabs(dd) / beta;
                sg = 1.0 / (beta * q * q);
            }
            const double lspike = (double)*ls + dl;
            const double lvNew = lvf;
            const double lsTot = (d.vReset - u) * lvNew + lpA * d.adaptInc + lspike;
            const double lt = lsTot * (1.0 - sd) * sg;
            const double lu = lvNew * (1.0 - s) + lt;
            const double lthr = -lt;
            // through sd: ds/d(sd) = 1 - st, d(sd)/d(rate) = dt/1000
            if (gDr) *gDr = (float)((double)*gDr + lsTot * (1.0 - st) * d.dtS);
            pA += lpA * s;
            pV += lthr;
            if (d.thrMask[i]) pK += lthr;
            ladf = (float)(lA + lthr * d.dA);
            lvf = (float)(lu * (1.0 + d.kM * (-1.0 - gEv * d.cE + gIv * d.cI)));
            lgEf = (float)(lE + lu * d.kM * (d.eExc - vv) * d.cE);
            lgIf = (float)(lI + lu * d.kM * (vv - d.eInh) * d.cI);
            if (bias != 0.0) pL += lu * d.kM;
            pEI += lu * d.kM * gIv * d.cI * (-1.0 + (vv - d.eInh) * d.cI);
            *ls = 0.f;
        }
        if (d.aClip > 0) {
            d.lv[i] = clip_to(lvf, d.aClip); d.lgE[i] = clip_to(lgEf, d.aClip); d.lgI[i] = clip_to(lgIf, d.aClip);
            d.lad[i] = clip_to(ladf, d.aClip); d.lres[i] = clip_to(lresf, d.aClip);
        } else {
            d.lv[i] = lvf; d.lgE[i] = lgEf; d.lgI[i] = lgIf; d.lad[i] = ladf; d.lres[i] = lresf;
        }
    }
    double p[5] = {pA, pV, pK, pL, pEI};
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < 5; c++) {
        const double x = wave_sum(p[c]);
        if (lane == 0) red[c][wave] = x;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        double s = 0.0;
        for (int w = 0; w < DIFF_BLOCK / 64; w++) s += red[threadIdx.x][w];
        d.blockAcc[(int64_t)blockIdx.x * 5 + threadIdx.x] += s;
    }
}

// Reversed delivery of step t (src/lifdiff.js:635-666) plus the per-edge gradient.
__device__ __forceinline__ void dx_barrival(const DiffDev &d, int32_t t, int32_t resFromT) {
    if (t < d.D) return;
    const int lane = threadIdx.x & 63;
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int64_t a0 = d.off[t - d.D], a1 = d.off[t - d.D + 1];
    const int64_t resBase = d.off[resFromT];
    const int32_t emitSlot = (t - d.D) % d.S;
    for (int64_t k = a0 + wave; k < a1; k += nwaves) {
        const int32_t pre = (int32_t)((uint32_t)d.idx[k] & ~FORCED_BIT);
        const double sgn = d.sign[pre];
        if (sgn == 0.0) continue;
        const double amp = d.amp[k];
        const double resAtv = d.resAt[k - resBase];
        const double gain = d.expLogGain[pre] * (double)d.outScale[pre];
        const double base = sgn * d.wSyn * gain * resAtv;
        const double eff = base * amp;
        const uint32_t ra = d.indptr[pre], re = d.indptr[pre + 1];
        double lRaw = 0.0, dInh = 0.0;
        if (eff > 0) {
            for (uint32_t j = ra + lane; j < re; j += 64) {
                const uint32_t q = d.indices[j];
                const double c = d.counts[j];
                const double cw = c * (d.edgeGain ? (double)d.edgeGain[j] : 1.0);
                const double lq = d.lgE[q], inS = d.inScale[q];
                lRaw += lq * cw * inS;
                atomicAdd(&d.gInScale[q], (float)(lq * cw * eff));
                d.gEdge[j] = (float)((double)d.gEdge[j] + lq * c * inS * eff);
            }
        } else {
            const double ig = d.inhGain;
            for (uint32_t j = ra + lane; j < re; j += 64) {
                const uint32_t q = d.indices[j];
                const double c = d.counts[j];
                const double cw = c * (d.edgeGain ? (double)d.edgeGain[j] : 1.0);
                const double lq = d.lgI[q], inS = d.inScale[q];
                lRaw += lq * cw * inS * ig;
                atomicAdd(&d.gInScale[q], (float)(lq * cw * ig * eff));
                dInh += lq * cw * inS * eff;
                d.gEdge[j] = (float)((double)d.gEdge[j] + lq * c * inS * ig * eff);
            }
        }
        lRaw = wave_sum(lRaw);
        dInh = wave_sum(dInh);
        if (lane == 0) {
            d.accW[pre] += lRaw * sgn * gain * resAtv * amp;
            d.accInh[pre] += dInh;
            d.gLogGain[pre] = (float)((double)d.gLogGain[pre] + lRaw * eff);
            float *tg = d.ring + (int64_t)emitSlot * d.N + pre;
            *tg = clip_to((double)*tg + lRaw * base, d.aClip);
            const double lresAfter = d.lres[pre];
            d.accDepU[pre] += -lresAfter * resAtv;
            d.lres[pre] = clip_to(lresAfter * (1.0 - d.depU) + lRaw * sgn * d.wSyn * gain * amp, d.aClip);
        }
    }
}

// ------------------------------------------------------------------ kernels
//
// Workgroups of 1024 threads that deliver in the fused forward launch (the
// same number of waves as the 256 x 256 delivery launch).
constexpr int DELIVER_WG = 64;
//
// Each step function runs as a kernel over one network (kd_*, the network by
// value) or over several networks in lockstep (kdb_*, blockIdx.y = network,
// descriptions in device memory; network y takes part while t < tEnd[y]).

#define FB_DIFF_KERNELS(NAME, BOUNDS, ARGS, CALL)                                                 \
    __global__ void __launch_bounds__(BOUNDS) kd_##NAME(DiffDev d ARGS) { CALL; }                 \
    __global__ void __launch_bounds__(BOUNDS) kdb_##NAME(const DiffDev *ds, const int32_t *tEnd,  \
                                                         int32_t tAt ARGS) {                      \
        if (tAt >= tEnd[blockIdx.y]) return;                                                     \
        const DiffDev d = ds[blockIdx.y];                                                        \
        CALL;                                                                                    \
    }

#define FB_COMMA ,
FB_DIFF_KERNELS(deliver_f, 256, FB_COMMA int32_t t FB_COMMA int32_t rf, (dx_deliver<false>(d, t, rf, gridDim.x)))
// Delivery and the drive pre-pass of one forward step in one launch: they
// touch disjoint state when the delay is at least two steps (delivery reads
// lists of steps t-D and t-D+1 and the sources' res; the pre-pass reads the
// driven neurons' refr and appends step t's list).  The last workgroup runs the
// pre-pass (at most 2048 driven neurons), the others deliver.
FB_DIFF_KERNELS(deliver_driven_f, 1024, FB_COMMA int32_t t,
                (blockIdx.x == gridDim.x - 1 ? dx_driven_fast(d, t) : dx_deliver<false>(d, t, 0, gridDim.x - 1)))
FB_DIFF_KERNELS(driven_f, 1024, FB_COMMA int32_t t,
                (d.nDriven <= 2048 ? dx_driven_fast(d, t) : dx_driven(d, t)))
// Replay: delivery and the forced spikes from the tape in one launch (disjoint
// state); the last workgroup applies the forced spikes.
FB_DIFF_KERNELS(deliver_forced_r, 256, FB_COMMA int32_t t FB_COMMA int32_t rf,
                (blockIdx.x == gridDim.x - 1 ? dx_forced_replay(d, t) : dx_deliver<true>(d, t, rf, gridDim.x - 1)))
FB_DIFF_KERNELS(long_s, 256, , (dx_long(d)))
FB_DIFF_KERNELS(membrane_f, DIFF_BLOCK, FB_COMMA int32_t t FB_COMMA int32_t r, (dx_membrane<false>(d, t, r)))
FB_DIFF_KERNELS(membrane_r, DIFF_BLOCK, FB_COMMA int32_t t FB_COMMA int32_t r, (dx_membrane<true>(d, t, r)))
FB_DIFF_KERNELS(restore, 256, FB_COMMA int32_t seg, (dx_restore(d, seg)))
FB_DIFF_KERNELS(save, 256, FB_COMMA int32_t seg, (dx_save(d, seg)))
FB_DIFF_KERNELS(bclear, 256, FB_COMMA int32_t ringToo, (dx_bclear(d, ringToo)))
FB_DIFF_KERNELS(bneuron, DIFF_BLOCK, FB_COMMA int32_t t FB_COMMA int32_t r, (dx_bneuron(d, t, r)))
FB_DIFF_KERNELS(barrival, 256, FB_COMMA int32_t t FB_COMMA int32_t rf, (dx_barrival(d, t, rf)))
#undef FB_COMMA
#undef FB_DIFF_KERNELS

// Final reduction into the nine scalar gradients (src/lifdiff.js:667-674), in
// a fixed order: one workgroup.
__global__ void __launch_bounds__(1024) kd_breduce(DiffDev d, int32_t nblk, double *out) {
    __shared__ double red[9][16];
    double p[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int32_t b = threadIdx.x; b < nblk; b += 1024) {
        p[7] += d.blockAcc[(int64_t)b * 5 + 0];   // adaptInc
        p[4] += d.blockAcc[(int64_t)b * 5 + 1];   // vThresh
        p[5] += d.blockAcc[(int64_t)b * 5 + 2];   // kcThreshold
        p[6] += d.blockAcc[(int64_t)b * 5 + 3];   // laminaBias
        p[3] += d.blockAcc[(int64_t)b * 5 + 4];   // eInh
    }
    for (int32_t i = threadIdx.x; i < d.N; i += 1024) {
        p[0] += d.accW[i];
        p[8] += d.accDepU[i];
        p[2] += d.accInh[i];
        const double sl = d.sizeLog[i];
        const double raw = exp(-d.sizeAlpha * sl);
        if (raw < d.boostCap && raw > d.minScale) p[1] += (double)d.gInScale[i] * (-sl) * raw;
    }
#pragma unroll
    for (int c = 0; c < 9; c++) {
        const double x = wave_sum(p[c]);
        if ((threadIdx.x & 63) == 0) red[c][threadIdx.x >> 6] = x;
    }
    __syncthreads();
    if (threadIdx.x < 9) {
        double s = 0.0;
        for (int w = 0; w < 16; w++) s += red[threadIdx.x][w];
        out[threadIdx.x] = s;
    }
}

// Non-finite gradients become zero and are counted (src/lifdiff.js:676-685).
__global__ void kd_sanitize(DiffDev d, int32_t *counts) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < d.N; i += stride)
        if (!isfinite(d.gLogGain[i])) { d.gLogGain[i] = 0.f; atomicAdd(&counts[1], 1); }
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < d.E; j += stride)
        if (!isfinite(d.gEdge[j])) { d.gEdge[j] = 0.f; atomicAdd(&counts[2], 1); }
}

}  // namespace fb

// ====================================================================== host

using namespace fb;

struct fb_diff {
    fb_diff_params p{};
    DiffDev d{};
    int32_t N = 0;
    int64_t E = 0;
    hipStream_t stream = nullptr, ownStream = nullptr;
    std::vector<void *> allocs;
    // host copies of the inputs
    std::vector<float> in[FB_D_NINPUTS];
    std::vector<int32_t> drivenList;
    std::vector<float> dL;
    std::vector<int32_t> readout;
    std::vector<float> target;
    double tauR = 20.0;
    int32_t lossFrom = 0;
    // device buffers whose size depends on the run
    float *ckpt = nullptr;            // [ncheckpoints][6N + 1]
    int32_t nck = 0;
    int64_t tapeCap = 0;
    int32_t tapeSteps = 0;
    int32_t steps = 0;                // of the recorded forward
    bool recorded = false;
    int64_t perStepCap = 0;           // tape entries allowed per step on average
    double *scal = nullptr;           // [0] flat loss, [1..9] params, [16..] readout parts
    double *readPart = nullptr;
    int32_t *clampDev = nullptr;
    double lastLoss = 0.0;
    double fwdMs = 0.0, bwdMs = 0.0, replayLaunches = 0.0;
    int32_t nblk = 0;
    std::vector<void *> runAllocs;    // freed and reallocated by forward()
    // the backward pass replays segment k-1 on its own stream while the
    // reverse sweep of segment k runs, so it keeps two arenas (by segment parity)
    float *arena2 = nullptr;
    uint8_t *mode2 = nullptr;
    hipStream_t replayStream = nullptr;
    hipEvent_t replayDone[2] = {nullptr, nullptr}, sweepDone[2] = {nullptr, nullptr}, fork = nullptr;
    int64_t arenaFloats = 0;          // floats of one arena (the modes are K N bytes more)
    // the drive as an input: schedule (host copy), per-epoch driven lists of a run
    std::vector<int32_t> drvIdx;
    std::vector<float> drvSched;
    std::vector<int32_t> epOff, epCnt;   // per epoch of the run: its driven list in epLists
    int32_t *epLists = nullptr;          // run allocation
    int32_t nEpPer = 0;
    void *schedAlloc = nullptr;          // drvSlot, drvSched, gDrive
    float *dPerAlloc = nullptr;
    double *lossAccAlloc = nullptr;
    bool thrMaskAny = false;             // exact32 refuses a threshold mask
    bool inScaleGiven = false;           // opts.inScale: in[FB_D_INSCALE_IN] taken verbatim
    float *betaAlloc = nullptr;          // per-neuron surrogateBeta (d.betaArr when set)
    float *x32Alloc = nullptr;           // exact32's W32 [E] and signW32 [N]
    void *trainAlloc = nullptr;          // the spike train: trainPtr [steps + 1], trainIdx
};

#define FB_REQUIRE_D(cond, code, ...)    \
    do {                                 \
        if (!(cond)) {                   \
            set_error(__VA_ARGS__);      \
            return code;                 \
        }                                \
    } while (0)

template <class T>
static fb_status dalloc(std::vector<void *> &list, T **p, size_t count) {
    void *q = nullptr;
    FB_HIP(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
    list.push_back(q);
    *p = (T *)q;
    return FB_OK;
}

static double js_round_d(double x) {
    const double f = std::floor(x);
    return (x - f >= 0.5) ? f + 1.0 : f;
}

static void derive(fb_diff *f) {
    const fb_diff_params &p = f->p;
    DiffDev &d = f->d;
    const double r = js_round_d(p.delay / p.dt);
    d.S = (int32_t)(r < 1 ? 1 : r) + 1;
    d.D = d.S - 1;
    d.K = p.checkpoint > 0 ? p.checkpoint : 16;
    d.soft = p.soft ? 1 : 0;
    d.dt = p.dt; d.vRest = p.vRest; d.vThresh = p.vThresh; d.vReset = p.vReset;
    d.wSyn = p.wSyn; d.adaptInc = p.adaptInc; d.depU = p.depU; d.eExc = p.eExc; d.eInh = p.eInh;
    d.inhGain = p.inhGain; d.kc = p.kcThreshold; d.lamina = p.laminaBias; d.beta = p.surrogateBeta;
    d.aClip = p.adjClip;
    d.dE = std::exp(-p.dt / p.tauSyn);
    d.dA = std::exp(-p.dt / p.adaptTau);
    d.kRec = p.dt / p.depTau;
    d.kM = p.dt / p.tauM;
    d.dtS = p.dt / 1000.0;
    d.cE = 1.0 / (p.eExc - p.vRest);
    d.cI = 1.0 / (p.vRest - p.eInh);
    d.emitThr = p.soft ? 1e-4 : 0.5;
    d.sizeAlpha = p.sizeAlpha;
    d.boostCap = p.boostCap;
    d.minScale = 1.0 / p.maxSizeScale;
    d.forcedRefr = (float)(p.tRef + p.dt);
    d.tRefF = (float)p.tRef;
    // lif.c reads every scalar from its binary32 header (src/lifdiff.js:231-240)
    d.exact32 = p.exact32 ? 1 : 0;
    d.xvRest = (float)p.vRest; d.xeExc = (float)p.eExc; d.xeInh = (float)p.eInh; d.xvThresh = (float)p.vThresh;
    d.xvReset = (float)p.vReset; d.xtRef = (float)p.tRef; d.xadaptInc = (float)p.adaptInc; d.xdt = (float)p.dt;
    d.xdepU = (float)p.depU; d.xlamina = (float)p.laminaBias;
    d.xdE = (float)d.dE; d.xdA = (float)d.dA; d.xkRec = (float)d.kRec; d.xkM = (float)d.kM; d.xdtS = (float)d.dtS;
    d.xcE = (float)d.cE; d.xcI = (float)d.cI;
    d.xforcedRefr = d.xtRef + d.xdt;
}

// inScale and exp(logGain), formed on the host in f64 (src/lifdiff.js:158-164, 278)
static fb_status upload_derived(fb_diff *f) {
    const int32_t N = f->N;
    std::vector<float> inS((size_t)N);
    std::vector<double> eg((size_t)N);
    const fb_diff_params &p = f->p;
    for (int32_t i = 0; i < N; i++) {
        if (f->inScaleGiven) {
            inS[i] = f->in[FB_D_INSCALE_IN][i];            // taken verbatim (src/lifdiff.js:160)
        } else {
            const double s = std::exp(-p.sizeAlpha * (double)f->in[FB_D_SIZELOG][i]);
            const double m = std::max(1.0 / p.maxSizeScale, s);
            inS[i] = (float)std::min(p.boostCap, m);
        }
        eg[i] = std::exp((double)f->in[FB_D_LOGGAIN][i]);
    }
    FB_HIP(hipMemcpy((void *)f->d.inScale, inS.data(), sizeof(float) * N, hipMemcpyHostToDevice));
    FB_HIP(hipMemcpy((void *)f->d.expLogGain, eg.data(), sizeof(double) * N, hipMemcpyHostToDevice));
    return FB_OK;
}

extern "C" fb_status fb_diff_create(int32_t N, int64_t E, const uint32_t *indptr, const uint32_t *indices,
                                    const float *counts, const float *sign, const fb_diff_params *p,
                                    fb_diff **out) {
    FB_REQUIRE_D(indptr && sign && p && out && (E == 0 || (indices && counts)), FB_ERR_ARG,
                 "fb_diff_create: null argument");
    FB_REQUIRE_D(N > 0 && E >= 0 && E < (1ll << 31), FB_ERR_ARG, "fb_diff_create: bad sizes N=%d E=%lld", N,
                 (long long)E);
    FB_REQUIRE_D(indptr[0] == 0 && (int64_t)indptr[N] == E, FB_ERR_ARG, "fb_diff_create: indptr must run 0..E");
    for (int32_t i = 0; i < N; i++)
        FB_REQUIRE_D(indptr[i] <= indptr[i + 1], FB_ERR_ARG, "fb_diff_create: indptr decreases at %d", i);
    for (int64_t j = 0; j < E; j++)
        FB_REQUIRE_D(indices[j] < (uint32_t)N, FB_ERR_ARG, "fb_diff_create: target %u out of range", indices[j]);
    FB_REQUIRE_D(p->dt > 0 && p->tauM > 0 && p->tauSyn > 0 && p->adaptTau > 0 && p->depTau > 0 && p->surrogateBeta > 0,
                 FB_ERR_ARG, "fb_diff_create: time constants and surrogateBeta must be positive");
    fb_diff *f = new fb_diff();
    f->p = *p;
    f->N = N;
    f->E = E;
    DiffDev &d = f->d;
    d.N = N;
    d.E = E;
    derive(f);
    f->nblk = (N + DIFF_BLOCK - 1) / DIFF_BLOCK;
    FB_HIP(hipStreamCreateWithFlags(&f->ownStream, hipStreamNonBlocking));
    f->stream = f->ownStream;
    for (int k = 0; k < FB_D_NINPUTS; k++) f->in[k].assign((size_t)N, k == FB_D_OUTSCALE ? 1.f : 0.f);
    f->dL.assign((size_t)N, 0.f);
    auto &A = f->allocs;
    uint32_t *ip, *ix;
    float *cn, *sg, *os, *to, *ins, *dr, *sl, *dls, *state;
    double *elg, *accs;
    uint8_t *masks, *forced;
    int32_t *drv, *dpos, *rid, *dflag, *cnt, *chain, *ncoll, *err, *longList, *nLong;
    double *add;
    uint64_t *key;
    DiffColl *coll;
    unsigned long long *top, *events;
    uint32_t *rng;
    fb_status st;
    const size_t n = (size_t)N;
    if ((st = dalloc(A, &ip, n + 1)) || (st = dalloc(A, &ix, (size_t)E)) || (st = dalloc(A, &cn, (size_t)E)) ||
        (st = dalloc(A, &sg, n)) || (st = dalloc(A, &os, n)) || (st = dalloc(A, &to, n)) || (st = dalloc(A, &ins, n)) ||
        (st = dalloc(A, &dr, n)) || (st = dalloc(A, &sl, n)) || (st = dalloc(A, &dls, n)) ||
        (st = dalloc(A, &elg, n)) || (st = dalloc(A, &masks, 2 * n)) || (st = dalloc(A, &drv, n)) ||
        (st = dalloc(A, &dpos, n)) || (st = dalloc(A, &rid, n)) || (st = dalloc(A, &dflag, n)) ||
        (st = dalloc(A, &state, 8 * n)) || (st = dalloc(A, &forced, n)) || (st = dalloc(A, &rng, 1)) ||
        (st = dalloc(A, &cnt, 2 * n)) || (st = dalloc(A, &chain, 2 * n)) || (st = dalloc(A, &add, 2 * n)) ||
        (st = dalloc(A, &key, 2 * n)) || (st = dalloc(A, &ncoll, 2)) || (st = dalloc(A, &err, 1)) ||
        (st = dalloc(A, &longList, 2 * n)) || (st = dalloc(A, &nLong, 1)) ||
        (st = dalloc(A, &top, 1)) || (st = dalloc(A, &events, 2)) || (st = dalloc(A, &accs, 3 * n + 5 * (size_t)f->nblk)) ||
        (st = dalloc(A, &f->scal, 16)) || (st = dalloc(A, &f->clampDev, 4)))
        return st;
    float *adj;
    if ((st = dalloc(A, &adj, (size_t)(5 + d.S + 3) * n + (size_t)E))) return st;
    d.collHalf = std::max<int64_t>(E, 1024);
    if ((st = dalloc(A, &coll, 2 * (size_t)d.collHalf))) return st;
    FB_HIP(hipMemcpy(ip, indptr, sizeof(uint32_t) * (n + 1), hipMemcpyHostToDevice));
    if (E) {
        FB_HIP(hipMemcpy(ix, indices, sizeof(uint32_t) * E, hipMemcpyHostToDevice));
        FB_HIP(hipMemcpy(cn, counts, sizeof(float) * E, hipMemcpyHostToDevice));
    }
    FB_HIP(hipMemcpy(sg, sign, sizeof(float) * n, hipMemcpyHostToDevice));
    d.indptr = ip; d.indices = ix; d.counts = cn; d.edgeGain = nullptr; d.sign = sg;
    d.outScale = os; d.thrOffset = to; d.inScale = ins; d.drive = dr; d.sizeLog = sl; d.dLdSpike = dls;
    d.expLogGain = elg; d.thrMask = masks; d.biasMask = masks + n;
    d.driven = drv; d.drvPos = dpos; d.readIdx = rid; d.dflag = dflag;
    d.v = state; d.gE = state + n; d.gI = state + 2 * n; d.adapt = state + 3 * n; d.res = state + 4 * n;
    d.refr = state + 5 * n; d.spikeCount = state + 6 * n; d.scored = state + 7 * n;
    d.forced = forced; d.rng = rng; d.cnt = cnt; d.chain = chain; d.add = add; d.key = key; d.coll = coll;
    d.ncoll = ncoll; d.err = err; d.top = top; d.events = events; d.longList = longList; d.nLong = nLong;
    d.accW = accs; d.accDepU = accs + n; d.accInh = accs + 2 * n; d.blockAcc = accs + 3 * n;
    d.lv = adj; d.lgE = adj + n; d.lgI = adj + 2 * n; d.lad = adj + 3 * n; d.lres = adj + 4 * n;
    d.ring = adj + 5 * n; d.gLogGain = adj + (5 + d.S) * n; d.gInScale = adj + (6 + d.S) * n;
    d.gEdge = adj + (8 + d.S) * n;
    d.nDriven = 0; d.R = 0; d.lossFrom = 0;
    uint32_t *jump = nullptr, *leap = nullptr;
    if ((st = leap_tables(&jump, &leap))) return st;
    d.leap = leap;
    std::vector<float> ones(n, 1.f), zeros(n, 0.f);
    std::vector<int32_t> neg(n, -1);
    FB_HIP(hipMemcpy(os, ones.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    for (float *z : {to, dr, sl, dls}) FB_HIP(hipMemcpy(z, zeros.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    FB_HIP(hipMemset(masks, 0, 2 * n));
    FB_HIP(hipMemcpy(dpos, neg.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    FB_HIP(hipMemcpy(rid, neg.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    FB_HIP(hipMemset(events, 0, 2 * sizeof(unsigned long long)));
    if ((st = upload_derived(f))) return st;
    f->perStepCap = std::max<int64_t>(4096, N / 4);
    *out = f;
    return FB_OK;
}

static void free_run(fb_diff *f) {
    for (void *q : f->runAllocs) (void)hipFree(q);
    f->runAllocs.clear();
    f->ckpt = nullptr;
    f->d.off = nullptr; f->d.idx = nullptr; f->d.amp = nullptr; f->d.readS = nullptr;
    f->d.aV = nullptr; f->d.rate = nullptr; f->d.dLds = nullptr;
    f->arena2 = nullptr; f->mode2 = nullptr;
    f->epLists = nullptr;
    f->recorded = false;
}

extern "C" fb_status fb_diff_destroy(fb_diff *f) {
    FB_REQUIRE_D(f, FB_ERR_ARG, "fb_diff_destroy: null handle");
    (void)hipStreamSynchronize(f->stream);
    free_run(f);
    for (void *q : f->allocs) (void)hipFree(q);
    if (f->d.edgeGain) (void)hipFree((void *)f->d.edgeGain);
    (void)hipFree((void *)f->d.target);
    (void)hipFree(f->readPart);
    if (f->schedAlloc) (void)hipFree(f->schedAlloc);
    if (f->dPerAlloc) (void)hipFree(f->dPerAlloc);
    if (f->lossAccAlloc) (void)hipFree(f->lossAccAlloc);
    if (f->betaAlloc) (void)hipFree(f->betaAlloc);
    if (f->x32Alloc) (void)hipFree(f->x32Alloc);
    if (f->trainAlloc) (void)hipFree(f->trainAlloc);
    (void)hipStreamDestroy(f->ownStream);
    if (f->replayStream) {
        (void)hipStreamDestroy(f->replayStream);
        for (int p = 0; p < 2; p++) {
            (void)hipEventDestroy(f->replayDone[p]);
            (void)hipEventDestroy(f->sweepDone[p]);
        }
        (void)hipEventDestroy(f->fork);
    }
    delete f;
    return FB_OK;
}

extern "C" fb_status fb_diff_set_params(fb_diff *f, const fb_diff_params *p) {
    FB_REQUIRE_D(f && p, FB_ERR_ARG, "fb_diff_set_params: null argument");
    FB_REQUIRE_D(p->dt > 0 && p->tauM > 0 && p->tauSyn > 0 && p->adaptTau > 0 && p->depTau > 0 && p->surrogateBeta > 0,
                 FB_ERR_ARG, "fb_diff_set_params: time constants and surrogateBeta must be positive");
    const int32_t S0 = f->d.S;
    fb_diff_params old = f->p;
    f->p = *p;
    derive(f);
    if (f->d.S != S0) {
        f->p = old;
        derive(f);
        set_error("fb_diff_set_params: delay / dt may not change the delay-line length");
        return FB_ERR_ARG;
    }
    f->recorded = false;
    return upload_derived(f);
}

extern "C" fb_status fb_diff_get_params(fb_diff *f, fb_diff_params *p) {
    FB_REQUIRE_D(f && p, FB_ERR_ARG, "fb_diff_get_params: null argument");
    *p = f->p;
    return FB_OK;
}

extern "C" fb_status fb_diff_set_input(fb_diff *f, int32_t which, const float *values) {
    FB_REQUIRE_D(f, FB_ERR_ARG, "fb_diff_set_input: null handle");
    FB_REQUIRE_D(which >= 0 && which < FB_D_NINPUTS, FB_ERR_ARG, "fb_diff_set_input: unknown input %d", which);
    FB_REQUIRE_D(values || which == FB_D_INSCALE_IN || which == FB_D_BETA, FB_ERR_ARG,
                 "fb_diff_set_input: null values (only inScale and surrogateBeta can be cleared)");
    const int32_t N = f->N;
    FB_HIP(hipStreamSynchronize(f->stream));
    if (which == FB_D_BETA) {
        if (values) {
            for (int32_t i = 0; i < N; i++)
                FB_REQUIRE_D(values[i] > 0, FB_ERR_ARG, "fb_diff_set_input: surrogateBeta[%d] must be positive", i);
            if (!f->betaAlloc) FB_HIP(hipMalloc(&f->betaAlloc, sizeof(float) * (size_t)N));
            FB_HIP(hipMemcpy(f->betaAlloc, values, sizeof(float) * N, hipMemcpyHostToDevice));
        }
        f->d.betaArr = values ? f->betaAlloc : nullptr;
        f->recorded = false;
        return FB_OK;
    }
    if (which == FB_D_INSCALE_IN) {
        f->inScaleGiven = values != nullptr;
        if (values) f->in[which].assign(values, values + N);
        f->recorded = false;
        return upload_derived(f);
    }
    f->in[which].assign(values, values + N);
    float *dst = nullptr;
    switch (which) {
        case FB_D_OUTSCALE: dst = (float *)f->d.outScale; break;
        case FB_D_THROFFSET: dst = (float *)f->d.thrOffset; break;
        case FB_D_SIZELOG: dst = (float *)f->d.sizeLog; break;
        case FB_D_DRIVE: dst = (float *)f->d.drive; break;
        default: break;
    }
    if (dst) FB_HIP(hipMemcpy(dst, values, sizeof(float) * N, hipMemcpyHostToDevice));
    f->recorded = false;
    return upload_derived(f);
}

extern "C" fb_status fb_diff_set_edge_gain(fb_diff *f, const float *edgeGain) {
    FB_REQUIRE_D(f, FB_ERR_ARG, "fb_diff_set_edge_gain: null handle");
    FB_HIP(hipStreamSynchronize(f->stream));
    if (!edgeGain) {
        if (f->d.edgeGain) FB_HIP(hipFree((void *)f->d.edgeGain));
        f->d.edgeGain = nullptr;
    } else {
        if (!f->d.edgeGain) {
            void *q = nullptr;
            FB_HIP(hipMalloc(&q, sizeof(float) * std::max<int64_t>(1, f->E)));
            f->d.edgeGain = (const float *)q;
        }
        if (f->E) FB_HIP(hipMemcpy((void *)f->d.edgeGain, edgeGain, sizeof(float) * f->E, hipMemcpyHostToDevice));
    }
    f->recorded = false;
    return FB_OK;
}

extern "C" fb_status fb_diff_set_masks(fb_diff *f, const uint8_t *thrMask, const uint8_t *biasMask) {
    FB_REQUIRE_D(f, FB_ERR_ARG, "fb_diff_set_masks: null handle");
    FB_HIP(hipStreamSynchronize(f->stream));
    if (thrMask) {
        FB_HIP(hipMemcpy((void *)f->d.thrMask, thrMask, (size_t)f->N, hipMemcpyHostToDevice));
        f->thrMaskAny = std::any_of(thrMask, thrMask + f->N, [](uint8_t m) { return m != 0; });
    }
    if (biasMask) FB_HIP(hipMemcpy((void *)f->d.biasMask, biasMask, (size_t)f->N, hipMemcpyHostToDevice));
    f->recorded = false;
    return FB_OK;
}

// The driven set in insertion order (LIFDiff.drivenSet, src/lifdiff.js:203-207).
extern "C" fb_status fb_diff_set_driven(fb_diff *f, const int32_t *idx, int32_t n) {
    FB_REQUIRE_D(f && n >= 0 && (n == 0 || idx), FB_ERR_ARG, "fb_diff_set_driven: bad arguments");
    std::vector<int32_t> pos((size_t)f->N, -1);
    for (int32_t q = 0; q < n; q++) {
        FB_REQUIRE_D(idx[q] >= 0 && idx[q] < f->N, FB_ERR_ARG, "fb_diff_set_driven: index %d out of range", idx[q]);
        FB_REQUIRE_D(pos[idx[q]] < 0, FB_ERR_ARG, "fb_diff_set_driven: neuron %d listed twice", idx[q]);
        pos[idx[q]] = q;
    }
    FB_HIP(hipStreamSynchronize(f->stream));
    f->drivenList.assign(idx, idx + n);
    if (n) FB_HIP(hipMemcpy((void *)f->d.driven, idx, sizeof(int32_t) * n, hipMemcpyHostToDevice));
    FB_HIP(hipMemcpy((void *)f->d.drvPos, pos.data(), sizeof(int32_t) * f->N, hipMemcpyHostToDevice));
    f->d.nDriven = n;
    f->recorded = false;
    return FB_OK;
}

extern "C" fb_status fb_diff_set_loss(fb_diff *f, const float *dLdSpike, int32_t lossFrom) {
    FB_REQUIRE_D(f && lossFrom >= 0, FB_ERR_ARG, "fb_diff_set_loss: bad arguments");
    FB_HIP(hipStreamSynchronize(f->stream));
    if (dLdSpike) f->dL.assign(dLdSpike, dLdSpike + f->N);
    else f->dL.assign((size_t)f->N, 0.f);
    FB_HIP(hipMemcpy((void *)f->d.dLdSpike, f->dL.data(), sizeof(float) * f->N, hipMemcpyHostToDevice));
    f->d.lossFrom = lossFrom;
    f->lossFrom = lossFrom;
    return FB_OK;
}

extern "C" fb_status fb_diff_set_drive_schedule(fb_diff *f, const int32_t *idx, int32_t nD, int32_t epochSteps,
                                               const float *sched, int32_t nEpochs, int32_t driveSoft) {
    FB_REQUIRE_D(f && nD >= 0 && (nD == 0 || (idx && sched && nEpochs > 0 && epochSteps > 0)), FB_ERR_ARG,
                 "fb_diff_set_drive_schedule: bad arguments");
    std::vector<int32_t> slot((size_t)f->N, -1);
    for (int32_t k = 0; k < nD; k++) {
        FB_REQUIRE_D(idx[k] >= 0 && idx[k] < f->N, FB_ERR_ARG, "fb_diff_set_drive_schedule: index %d out of range",
                     idx[k]);
        FB_REQUIRE_D(slot[idx[k]] < 0, FB_ERR_ARG, "fb_diff_set_drive_schedule: neuron %d listed twice", idx[k]);
        slot[idx[k]] = k;
    }
    FB_HIP(hipStreamSynchronize(f->stream));
    if (f->schedAlloc) FB_HIP(hipFree(f->schedAlloc));
    f->schedAlloc = nullptr;
    f->d.drvSlot = nullptr; f->d.drvSched = nullptr; f->d.gDrive = nullptr;
    f->d.nDI = nD; f->d.nEp = nD ? nEpochs : 0; f->d.driveSoft = driveSoft ? 1 : 0;
    if (nD || f->d.dPer == nullptr) f->d.ep = nD ? epochSteps : (f->d.dPer ? f->d.ep : 1);
    f->drvIdx.assign(idx, idx + nD);
    f->drvSched.assign(sched, sched + (size_t)(nD ? nEpochs : 0) * nD);
    if (nD) {
        const size_t nv = (size_t)nEpochs * nD;
        FB_HIP(hipMalloc(&f->schedAlloc, sizeof(int32_t) * f->N + 2 * sizeof(float) * nv));
        int32_t *ds = (int32_t *)f->schedAlloc;
        float *sv = (float *)(ds + f->N);
        FB_HIP(hipMemcpy(ds, slot.data(), sizeof(int32_t) * f->N, hipMemcpyHostToDevice));
        FB_HIP(hipMemcpy(sv, sched, sizeof(float) * nv, hipMemcpyHostToDevice));
        FB_HIP(hipMemset(sv + nv, 0, sizeof(float) * nv));
        f->d.drvSlot = ds; f->d.drvSched = sv; f->d.gDrive = sv + nv;
    }
    f->recorded = false;
    return FB_OK;
}

extern "C" fb_status fb_diff_set_loss_per_epoch(fb_diff *f, const float *dPer, int32_t nEpochs, int32_t epochSteps) {
    FB_REQUIRE_D(f && (!dPer || (nEpochs > 0 && epochSteps > 0)), FB_ERR_ARG, "fb_diff_set_loss_per_epoch: bad arguments");
    FB_REQUIRE_D(!dPer || !f->d.nDI || epochSteps == f->d.ep, FB_ERR_ARG,
                 "fb_diff_set_loss_per_epoch: the epoch must be the drive schedule's (%d steps)", f->d.ep);
    FB_HIP(hipStreamSynchronize(f->stream));
    if (f->dPerAlloc) FB_HIP(hipFree(f->dPerAlloc));
    f->dPerAlloc = nullptr;
    f->d.dPer = nullptr;
    f->nEpPer = 0;
    if (dPer) {
        const size_t nv = (size_t)nEpochs * f->N;
        FB_HIP(hipMalloc(&f->dPerAlloc, sizeof(float) * nv));
        FB_HIP(hipMemcpy(f->dPerAlloc, dPer, sizeof(float) * nv, hipMemcpyHostToDevice));
        if (!f->lossAccAlloc) FB_HIP(hipMalloc(&f->lossAccAlloc, sizeof(double) * f->N));
        f->d.dPer = f->dPerAlloc;
        f->d.lossAcc = f->lossAccAlloc;
        f->d.ep = epochSteps;
        f->nEpPer = nEpochs;
    }
    // the tape does not depend on the weights (backward(tape, dLdSpike,
    // {dLdSpikePerEpoch}) takes them after the forward pass); only the loss
    // the forward pass reported does
    return FB_OK;
}

extern "C" fb_status fb_diff_set_spike_train(fb_diff *f, int32_t steps, const int64_t *ptr, const int32_t *idx) {
    FB_REQUIRE_D(f && steps >= 0 && (steps == 0 || ptr), FB_ERR_ARG, "fb_diff_set_spike_train: bad arguments");
    if (steps > 0) {
        FB_REQUIRE_D(ptr[0] == 0, FB_ERR_ARG, "fb_diff_set_spike_train: ptr[0] must be 0");
        for (int32_t t = 0; t < steps; t++)
            FB_REQUIRE_D(ptr[t + 1] >= ptr[t], FB_ERR_ARG, "fb_diff_set_spike_train: ptr must not decrease (step %d)", t);
        FB_REQUIRE_D(ptr[steps] == 0 || idx, FB_ERR_ARG, "fb_diff_set_spike_train: null idx");
        for (int64_t k = 0; k < ptr[steps]; k++)
            FB_REQUIRE_D(idx[k] >= 0 && idx[k] < f->N, FB_ERR_ARG, "fb_diff_set_spike_train: neuron %d out of range",
                         idx[k]);
    }
    FB_HIP(hipStreamSynchronize(f->stream));
    if (f->trainAlloc) FB_HIP(hipFree(f->trainAlloc));
    f->trainAlloc = nullptr;
    f->d.trainSteps = 0;
    f->d.trainPtr = nullptr;
    f->d.trainIdx = nullptr;
    if (steps > 0) {
        const int64_t nIdx = ptr[steps];
        const size_t ptrBytes = sizeof(int64_t) * ((size_t)steps + 1);
        FB_HIP(hipMalloc(&f->trainAlloc, ptrBytes + sizeof(int32_t) * (size_t)std::max<int64_t>(1, nIdx)));
        char *base = (char *)f->trainAlloc;
        FB_HIP(hipMemcpy(base, ptr, ptrBytes, hipMemcpyHostToDevice));
        if (nIdx) FB_HIP(hipMemcpy(base + ptrBytes, idx, sizeof(int32_t) * (size_t)nIdx, hipMemcpyHostToDevice));
        f->d.trainSteps = steps;
        f->d.trainPtr = (const int64_t *)base;
        f->d.trainIdx = (const int32_t *)(base + ptrBytes);
    }
    // a recorded tape keeps the train it ran with (its forced spikes are on
    // the tape, and the replay applies them from there)
    return FB_OK;
}

extern "C" fb_status fb_diff_drive_grad(fb_diff *f, float *gDrive) {
    FB_REQUIRE_D(f && gDrive, FB_ERR_ARG, "fb_diff_drive_grad: bad arguments");
    FB_REQUIRE_D(f->d.gDrive, FB_ERR_STATE, "fb_diff_drive_grad: no drive schedule");
    FB_HIP(hipStreamSynchronize(f->stream));
    const size_t nv = (size_t)f->d.nEp * f->d.nDI;
    FB_HIP(hipMemcpy(gDrive, f->d.gDrive, sizeof(float) * nv, hipMemcpyDeviceToHost));
    for (size_t q = 0; q < nv; q++)
        if (!std::isfinite(gDrive[q])) gDrive[q] = 0.f;   // as the reference sanitises (src/lifdiff.js:684)
    return FB_OK;
}

extern "C" fb_status fb_diff_set_readout(fb_diff *f, const int32_t *idx, int32_t R, const float *targetHz,
                                         double tauMs) {
    FB_REQUIRE_D(f && R >= 0 && (R == 0 || (idx && targetHz)) && tauMs > 0, FB_ERR_ARG,
                 "fb_diff_set_readout: bad arguments");
    std::vector<int32_t> rid((size_t)f->N, -1);
    for (int32_t k = 0; k < R; k++) {
        FB_REQUIRE_D(idx[k] >= 0 && idx[k] < f->N, FB_ERR_ARG, "fb_diff_set_readout: index %d out of range", idx[k]);
        FB_REQUIRE_D(rid[idx[k]] < 0, FB_ERR_ARG, "fb_diff_set_readout: neuron %d listed twice", idx[k]);
        rid[idx[k]] = k;
    }
    FB_HIP(hipStreamSynchronize(f->stream));
    FB_HIP(hipMemcpy((void *)f->d.readIdx, rid.data(), sizeof(int32_t) * f->N, hipMemcpyHostToDevice));
    if (f->d.target) FB_HIP(hipFree((void *)f->d.target));
    if (f->readPart) FB_HIP(hipFree(f->readPart));
    void *q = nullptr;
    FB_HIP(hipMalloc(&q, sizeof(float) * std::max(1, R)));
    f->d.target = (const float *)q;
    FB_HIP(hipMalloc(&f->readPart, sizeof(double) * std::max(1, R)));
    if (R) FB_HIP(hipMemcpy(q, targetHz, sizeof(float) * R, hipMemcpyHostToDevice));
    f->readout.assign(idx, idx + R);
    f->target.assign(targetHz, targetHz + R);
    f->tauR = tauMs;
    f->d.R = R;
    f->recorded = false;
    return FB_OK;
}

static int grid_of(int64_t n, int block, int cap = 4096) {
    return (int)std::max<int64_t>(1, std::min<int64_t>(cap, (n + block - 1) / block));
}

static fb_status check_err(fb_diff *f, const char *who) {
    int32_t err = 0;
    FB_HIP(hipMemcpyAsync(&err, f->d.err, sizeof(int32_t), hipMemcpyDeviceToHost, f->stream));
    FB_HIP(hipStreamSynchronize(f->stream));
    FB_REQUIRE_D(!(err & DERR_COLL), FB_ERR_OVERFLOW, "%s: arrival chain pool overflowed", who);
    FB_REQUIRE_D(!(err & DERR_TAPE), FB_ERR_OVERFLOW,
                 "%s: the spike tape overflowed (raise the per-step capacity with fb_diff_set_tape_capacity)", who);
    return FB_OK;
}

extern "C" fb_status fb_diff_set_tape_capacity(fb_diff *f, int64_t perStep) {
    FB_REQUIRE_D(f && perStep > 0, FB_ERR_ARG, "fb_diff_set_tape_capacity: bad arguments");
    f->perStepCap = perStep;
    return FB_OK;
}

// Buffers and initial state of a forward pass of `steps` steps, queued on st.
static fb_status fwd_prepare(fb_diff *f, int32_t steps, double seed, int32_t record, hipStream_t st) {
    FB_REQUIRE_D(f->d.nDriven <= 1 << 20, FB_ERR_UNSUPPORTED, "fb_diff_forward: driven set too large");
    DiffDev &d = f->d;
    const int32_t N = f->N, R = d.R;
    FB_HIP(hipStreamSynchronize(f->stream));
    free_run(f);
    fb_status s;
    const int64_t cap = std::max<int64_t>(1024, f->perStepCap * (int64_t)std::max(1, steps));
    if ((s = dalloc(f->runAllocs, &d.off, (size_t)steps + 1)) || (s = dalloc(f->runAllocs, &d.idx, (size_t)cap)) ||
        (s = dalloc(f->runAllocs, &d.amp, (size_t)cap)) ||
        (s = dalloc(f->runAllocs, &d.readS, (size_t)std::max(1, steps) * std::max(1, R))) ||
        (s = dalloc(f->runAllocs, &d.rate, (size_t)std::max(1, steps) * std::max(1, R))) ||
        (s = dalloc(f->runAllocs, &d.dLds, (size_t)std::max(1, steps) * std::max(1, R))))
        return s;
    d.cap = cap;
    if (R) FB_HIP(hipMemsetAsync(d.readS, 0, sizeof(float) * (size_t)steps * R, st));
    f->nck = record ? (steps + d.K - 1) / d.K + 1 : 0;
    if (record && (s = dalloc(f->runAllocs, &f->ckpt, (size_t)f->nck * (6 * (size_t)N + 1)))) return s;
    d.ckpt = f->ckpt;
    FB_REQUIRE_D(!d.nDI || (int64_t)d.nEp * d.ep >= steps, FB_ERR_ARG,
                 "fb_diff_forward: the drive schedule covers %d epochs of %d steps, fewer than %d steps", d.nEp, d.ep,
                 steps);
    FB_REQUIRE_D(!d.dPer || (int64_t)f->nEpPer * d.ep >= steps, FB_ERR_ARG,
                 "fb_diff_forward: the per-epoch spike weights cover fewer than %d steps", steps);
    if (d.dPer) FB_HIP(hipMemsetAsync(d.lossAcc, 0, sizeof(double) * N, st));
    // the driven set at each epoch: the inputs' set, then setDriveOne with the
    // epoch's rates (src/lifdiff.js:205-207): an added neuron goes to the end,
    // a removed one leaves its place
    f->epOff.clear(); f->epCnt.clear();
    if (d.nDI && !d.driveSoft) {
        const int32_t nE = (steps + d.ep - 1) / d.ep;
        std::vector<int32_t> cur(f->drivenList), all;
        std::vector<uint8_t> in((size_t)N, 0);
        for (int32_t i : cur) in[i] = 1;
        for (int32_t e = 0; e < nE; e++) {
            const int32_t ee = std::min(e, d.nEp - 1);
            for (int32_t k = 0; k < d.nDI; k++) {
                const int32_t i = f->drvIdx[k];
                const float rate = f->drvSched[(size_t)ee * d.nDI + k];
                // -0.0: a positive rate too small for binary32, which joins
                // the set with a zero drive
                if (rate > 0 || (rate == 0.f && std::signbit(rate))) {
                    if (!in[i]) { in[i] = 1; cur.push_back(i); }
                } else if (in[i]) {
                    in[i] = 0;
                    cur.erase(std::find(cur.begin(), cur.end(), i));
                }
            }
            f->epOff.push_back((int32_t)all.size());
            f->epCnt.push_back((int32_t)cur.size());
            all.insert(all.end(), cur.begin(), cur.end());
        }
        if ((s = dalloc(f->runAllocs, &f->epLists, std::max<size_t>(all.size(), 1)))) return s;
        if (!all.empty())
            FB_HIP(hipMemcpyAsync(f->epLists, all.data(), sizeof(int32_t) * all.size(), hipMemcpyHostToDevice, st));
        FB_HIP(hipStreamSynchronize(st));   // `all` leaves scope
    }
    if (d.exact32) {
        // lif.c sums vThresh + adapt + thr only (src/lifdiff.js:178-180)
        FB_REQUIRE_D(!f->thrMaskAny, FB_ERR_ARG,
                     "fb_diff_forward: exact32 needs the whole threshold field in thrOffset; thrMask must be empty");
        if (!f->x32Alloc) FB_HIP(hipMalloc(&f->x32Alloc, sizeof(float) * ((size_t)std::max<int64_t>(f->E, 1) + N)));
        d.W32 = f->x32Alloc;
        d.signW32 = f->x32Alloc + std::max<int64_t>(f->E, 1);
        hipLaunchKernelGGL(kd_bake_x32, dim3(grid_of((int64_t)N * 64, 256, 8192)), dim3(256), 0, st, d);
    }
    const uint32_t r0 = seed_hash(seed);
    FB_HIP(hipMemcpyAsync(d.rng, &r0, sizeof r0, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(kd_reset, dim3(grid_of(N, 256)), dim3(256), 0, st, d, (float)f->p.vRest);
    FB_HIP(hipGetLastError());
    return FB_OK;
}

// After the steps: the tape's end, the losses
static fb_status fwd_finish_launch(fb_diff *f, int32_t steps, hipStream_t st) {
    const DiffDev &d = f->d;
    const int32_t R = d.R;
    hipLaunchKernelGGL(kd_finish, dim3(1), dim3(64), 0, st, d, steps);
    hipLaunchKernelGGL(kd_loss_flat, dim3(1), dim3(1024), 0, st, d, f->scal);
    if (R && steps) {
        const double a = std::exp(-f->p.dt / f->tauR);
        hipLaunchKernelGGL(kd_readout, dim3((R + 255) / 256), dim3(256), 0, st, d, steps, a, f->readPart);
    }
    FB_HIP(hipGetLastError());
    return FB_OK;
}

// (after the stream is synchronised) the loss and the run's bookkeeping
static fb_status fwd_collect(fb_diff *f, int32_t steps, int32_t record, double *loss) {
    fb_status s;
    if ((s = check_err(f, "fb_diff_forward"))) return s;
    const int32_t R = f->d.R;
    double L = 0.0;
    FB_HIP(hipMemcpy(&L, f->scal, sizeof(double), hipMemcpyDeviceToHost));
    if (R && steps) {
        std::vector<double> parts((size_t)R);
        FB_HIP(hipMemcpy(parts.data(), f->readPart, sizeof(double) * R, hipMemcpyDeviceToHost));
        double Lr = 0.0;
        for (int32_t k = 0; k < R; k++) Lr += parts[k];
        L += Lr / ((double)steps * (double)R);
    }
    f->steps = steps;
    f->recorded = record != 0;
    f->lastLoss = L;
    if (loss) *loss = L;
    return FB_OK;
}

extern "C" fb_status fb_diff_forward(fb_diff *f, int32_t steps, double seed, int32_t record, double *loss) {
    FB_REQUIRE_D(f && steps >= 0, FB_ERR_ARG, "fb_diff_forward: bad arguments");
    DiffDev &d = f->d;
    const int32_t N = f->N;
    hipStream_t st = f->stream;
    fb_status s;
    if ((s = fwd_prepare(f, steps, seed, record, st))) return s;
    hipEvent_t e0, e1;
    FB_HIP(hipEventCreate(&e0));
    FB_HIP(hipEventCreate(&e1));
    FB_HIP(hipEventRecord(e0, st));
    const int gN = grid_of(N, 256);
    for (int32_t t = 0; t < steps; t++) {
        // the driven set of the step's epoch (none with the soft drive)
        DiffDev ds = d;
        if (d.driveSoft) {
            ds.nDriven = 0;
        } else if (!f->epCnt.empty()) {
            const int32_t e = t / d.ep;
            ds.driven = f->epLists + f->epOff[(size_t)e];
            ds.nDriven = f->epCnt[(size_t)e];
        }
        const bool fused = ds.D >= 2 && ds.nDriven <= 2048;
        if (record && t % d.K == 0) hipLaunchKernelGGL(kd_save, dim3(gN), dim3(256), 0, st, ds, t / d.K);
        if (fused) {
            hipLaunchKernelGGL(kd_deliver_driven_f, dim3(DELIVER_WG + 1), dim3(1024), 0, st, ds, t);
        } else {
            hipLaunchKernelGGL(kd_deliver_f, dim3(256), dim3(256), 0, st, ds, t, 0);
            hipLaunchKernelGGL(kd_driven_f, dim3(1), dim3(1024), 0, st, ds, t);
        }
        hipLaunchKernelGGL(kd_long_s, dim3(DIFF_LONG_WG), dim3(256), 0, st, ds);
        hipLaunchKernelGGL(kd_membrane_f, dim3(f->nblk), dim3(DIFF_BLOCK), 0, st, ds, t, 0);
    }
    FB_HIP(hipGetLastError());
    if ((s = fwd_finish_launch(f, steps, st))) return s;
    FB_HIP(hipEventRecord(e1, st));
    FB_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    FB_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    f->fwdMs = ms;
    return fwd_collect(f, steps, record, loss);
}

// Arena of one checkpoint segment and the cleared adjoint state, queued on st.
static fb_status bwd_prepare(fb_diff *f, hipStream_t st) {
    FB_REQUIRE_D(f->recorded, FB_ERR_STATE, "fb_diff_backward: no recorded forward pass (forward with record=1, "
                                            "after the last change of inputs)");
    DiffDev &d = f->d;
    const int32_t N = f->N, T = f->steps, K = d.K, D = d.D;
    fb_status s;
    if (!d.aV) {
        float *ar;
        uint8_t *md;
        int64_t resCap = 1;
        std::vector<int64_t> off((size_t)T + 1);
        FB_HIP(hipMemcpy(off.data(), d.off, sizeof(int64_t) * (T + 1), hipMemcpyDeviceToHost));
        for (int32_t from = 0; from < T; from += K) {
            const int32_t to = std::min(T, from + K);
            if (to > D) resCap = std::max<int64_t>(resCap, off[to - D] - off[std::max(from, D) - D]);
        }
        if ((s = dalloc(f->runAllocs, &ar, 5 * (size_t)K * N + (size_t)resCap)) ||
            (s = dalloc(f->runAllocs, &md, (size_t)K * N)))
            return s;
        d.aV = ar; d.aGE = ar + (size_t)K * N; d.aGI = ar + 2 * (size_t)K * N; d.aAd = ar + 3 * (size_t)K * N;
        d.aSp = ar + 4 * (size_t)K * N; d.resAt = ar + 5 * (size_t)K * N; d.aMode = md;
        f->arenaFloats = 5 * (int64_t)K * N + resCap;
    }
    hipLaunchKernelGGL(kd_bclear, dim3(grid_of(N, 256)), dim3(256), 0, st, d, 1);
    hipLaunchKernelGGL(kd_bzero, dim3(grid_of(std::max<int64_t>(f->E, N), 256)), dim3(256), 0, st, d, f->nblk);
    if (d.gDrive) FB_HIP(hipMemsetAsync(d.gDrive, 0, sizeof(float) * (size_t)d.nEp * d.nDI, st));
    FB_HIP(hipMemsetAsync(f->clampDev, 0, 4 * sizeof(int32_t), st));
    FB_HIP(hipGetLastError());
    return FB_OK;
}

static fb_status bwd_finish_launch(fb_diff *f, hipStream_t st) {
    const DiffDev &d = f->d;
    hipLaunchKernelGGL(kd_breduce, dim3(1), dim3(1024), 0, st, d, f->nblk, f->scal + 1);
    hipLaunchKernelGGL(kd_sanitize, dim3(grid_of(std::max<int64_t>(f->E, f->N), 256)), dim3(256), 0, st, d,
                       f->clampDev);
    FB_HIP(hipGetLastError());
    return FB_OK;
}

// The results of the last backward pass (any output may be NULL).
extern "C" fb_status fb_diff_grads(fb_diff *f, double *params, float *gLogGain, float *gEdge, float *gInScale,
                                   int32_t *clamped) {
    FB_REQUIRE_D(f, FB_ERR_ARG, "fb_diff_grads: null handle");
    FB_HIP(hipStreamSynchronize(f->stream));
    double g[9];
    int32_t cl[4];
    FB_HIP(hipMemcpy(g, f->scal + 1, sizeof g, hipMemcpyDeviceToHost));
    FB_HIP(hipMemcpy(cl, f->clampDev, sizeof cl, hipMemcpyDeviceToHost));
    cl[0] = 0;
    for (int q = 0; q < 9; q++)
        if (!std::isfinite(g[q])) { g[q] = 0.0; cl[0]++; }
    if (params) std::memcpy(params, g, sizeof g);
    if (clamped) std::memcpy(clamped, cl, 3 * sizeof(int32_t));
    if (gLogGain) FB_HIP(hipMemcpy(gLogGain, f->d.gLogGain, sizeof(float) * f->N, hipMemcpyDeviceToHost));
    if (gInScale) FB_HIP(hipMemcpy(gInScale, f->d.gInScale, sizeof(float) * f->N, hipMemcpyDeviceToHost));
    if (gEdge && f->E) FB_HIP(hipMemcpy(gEdge, f->d.gEdge, sizeof(float) * f->E, hipMemcpyDeviceToHost));
    return FB_OK;
}

// The arena of segment parity 1 (allocated with the run's buffers).
static fb_status alloc_arena2(fb_diff *f) {
    fb_status s;
    if (!f->arena2 &&
        ((s = dalloc(f->runAllocs, &f->arena2, (size_t)f->arenaFloats)) ||
         (s = dalloc(f->runAllocs, &f->mode2, (size_t)f->d.K * f->N))))
        return s;
    return FB_OK;
}

// The same description with the arena of segment parity 1.
static DiffDev with_arena2(const fb_diff *f) {
    DiffDev d = f->d;
    const int64_t KN = (int64_t)d.K * f->N;
    float *ar = f->arena2;
    d.aV = ar; d.aGE = ar + KN; d.aGI = ar + 2 * KN; d.aAd = ar + 3 * KN; d.aSp = ar + 4 * KN; d.resAt = ar + 5 * KN;
    d.aMode = f->mode2;
    return d;
}

extern "C" fb_status fb_diff_backward(fb_diff *f, int32_t truncate, double *params, float *gLogGain, float *gEdge,
                                      float *gInScale, int32_t *clamped) {
    FB_REQUIRE_D(f && truncate >= 0, FB_ERR_ARG, "fb_diff_backward: bad arguments");
    const int32_t N = f->N, T = f->steps, K = f->d.K, D = f->d.D;
    hipStream_t st = f->stream;
    fb_status s;
    hipEvent_t e0, e1;
    FB_HIP(hipEventCreate(&e0));
    FB_HIP(hipEventCreate(&e1));
    if ((s = bwd_prepare(f, st)) || (s = alloc_arena2(f))) return s;
    if (!f->replayStream) {
        FB_HIP(hipStreamCreateWithFlags(&f->replayStream, hipStreamNonBlocking));
        for (int p = 0; p < 2; p++) {
            FB_HIP(hipEventCreateWithFlags(&f->replayDone[p], hipEventDisableTiming));
            FB_HIP(hipEventCreateWithFlags(&f->sweepDone[p], hipEventDisableTiming));
        }
        FB_HIP(hipEventCreateWithFlags(&f->fork, hipEventDisableTiming));
    }
    const DiffDev dp[2] = {f->d, with_arena2(f)};   // by segment parity
    hipStream_t rs = f->replayStream;
    FB_HIP(hipEventRecord(e0, st));
    FB_HIP(hipEventRecord(f->fork, st));
    FB_HIP(hipStreamWaitEvent(rs, f->fork, 0));
    const int gN = grid_of(N, 256);
    const int32_t nseg = (T + K - 1) / K;
    // segment sg: restore its checkpoint and re-run it forward, filling the arena of its parity
    auto replay = [&](int32_t sg) {
        const DiffDev &d = dp[sg & 1];
        const int32_t from = sg * K, to = std::min(T, from + K);
        const int32_t resFromT = std::max(from, D) - D;
        hipLaunchKernelGGL(kd_restore, dim3(gN), dim3(256), 0, rs, d, sg);
        for (int32_t t = from; t < to; t++) {
            hipLaunchKernelGGL(kd_deliver_forced_r, dim3(257), dim3(256), 0, rs, d, t, resFromT);
            hipLaunchKernelGGL(kd_long_s, dim3(DIFF_LONG_WG), dim3(256), 0, rs, d);
            hipLaunchKernelGGL(kd_membrane_r, dim3(f->nblk), dim3(DIFF_BLOCK), 0, rs, d, t, t - from);
        }
    };
    if (nseg > 0) {
        replay(nseg - 1);
        FB_HIP(hipEventRecord(f->replayDone[(nseg - 1) & 1], rs));
    }
    for (int32_t sgi = nseg - 1; sgi >= 0; sgi--) {
        // the next replay overlaps this sweep; it reuses the arena of sweep sgi + 1
        if (sgi > 0) {
            if (sgi + 1 < nseg) FB_HIP(hipStreamWaitEvent(rs, f->sweepDone[(sgi + 1) & 1], 0));
            replay(sgi - 1);
            FB_HIP(hipEventRecord(f->replayDone[(sgi - 1) & 1], rs));
        }
        const DiffDev &d = dp[sgi & 1];
        const int32_t from = sgi * K, to = std::min(T, from + K);
        const int32_t resFromT = std::max(from, D) - D;
        FB_HIP(hipStreamWaitEvent(st, f->replayDone[sgi & 1], 0));
        for (int32_t t = to - 1; t >= from; t--) {
            if (truncate > 0 && (t + 1) % truncate == 0)
                hipLaunchKernelGGL(kd_bclear, dim3(gN), dim3(256), 0, st, d, 1);
            hipLaunchKernelGGL(kd_bneuron, dim3(f->nblk), dim3(DIFF_BLOCK), 0, st, d, t, t - from);
            hipLaunchKernelGGL(kd_barrival, dim3(256), dim3(256), 0, st, d, t, resFromT);
        }
        FB_HIP(hipEventRecord(f->sweepDone[sgi & 1], st));
        FB_HIP(hipGetLastError());
    }
    if ((s = bwd_finish_launch(f, st))) return s;
    FB_HIP(hipEventRecord(e1, st));
    if ((s = check_err(f, "fb_diff_backward"))) return s;
    float ms = 0.f;
    FB_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    f->bwdMs = ms;
    return fb_diff_grads(f, params, gLogGain, gEdge, gInScale, clamped);
}

// ------------------------------------------------------------------ several networks in lockstep

struct fb_diff_batch {
    std::vector<fb_diff *> m;
    DiffDev *ds = nullptr;              // device copies of the members' descriptions
    int32_t *tEnd = nullptr;            // per member: steps below this take part
    hipStream_t stream = nullptr;
    std::vector<int32_t> steps;
    double fwdMs = 0.0, bwdMs = 0.0;
    int32_t maxBlk = 0;
    int32_t *segEnds = nullptr;         // backward: every segment's per-member ends [nseg][B]
    int64_t segEndsCap = 0;
};

extern "C" fb_status fb_diff_batch_create(fb_diff *const *members, int32_t B, fb_diff_batch **out) {
    FB_REQUIRE_D(members && out && B >= 1, FB_ERR_ARG, "fb_diff_batch_create: bad arguments");
    for (int32_t b = 0; b < B; b++) {
        FB_REQUIRE_D(members[b], FB_ERR_ARG, "fb_diff_batch_create: null member %d", b);
        FB_REQUIRE_D(members[b]->d.S == members[0]->d.S && members[b]->d.K == members[0]->d.K, FB_ERR_ARG,
                     "fb_diff_batch_create: members differ in delay length or checkpoint interval");
        for (int32_t c = 0; c < b; c++)
            FB_REQUIRE_D(members[c] != members[b], FB_ERR_ARG, "fb_diff_batch_create: member %d listed twice", b);
    }
    fb_diff_batch *bt = new fb_diff_batch();
    bt->m.assign(members, members + B);
    FB_HIP(hipMalloc(&bt->ds, sizeof(DiffDev) * B));
    FB_HIP(hipMalloc(&bt->tEnd, sizeof(int32_t) * B));
    FB_HIP(hipStreamCreateWithFlags(&bt->stream, hipStreamNonBlocking));
    for (fb_diff *f : bt->m) bt->maxBlk = std::max(bt->maxBlk, f->nblk);
    *out = bt;
    return FB_OK;
}

extern "C" fb_status fb_diff_batch_destroy(fb_diff_batch *bt) {
    FB_REQUIRE_D(bt, FB_ERR_ARG, "fb_diff_batch_destroy: null batch");
    (void)hipStreamSynchronize(bt->stream);
    (void)hipFree(bt->ds);
    (void)hipFree(bt->tEnd);
    (void)hipStreamDestroy(bt->stream);
    (void)hipFree(bt->segEnds);
    delete bt;
    return FB_OK;
}

static fb_status batch_upload(fb_diff_batch *bt, const std::vector<int32_t> &tEnd) {
    std::vector<DiffDev> h;
    for (fb_diff *f : bt->m) h.push_back(f->d);
    FB_HIP(hipMemcpyAsync(bt->ds, h.data(), sizeof(DiffDev) * h.size(), hipMemcpyHostToDevice, bt->stream));
    return FB_OK;
}

static fb_status batch_ends(fb_diff_batch *bt, const std::vector<int32_t> &tEnd) {
    FB_HIP(hipMemcpyAsync(bt->tEnd, tEnd.data(), sizeof(int32_t) * tEnd.size(), hipMemcpyHostToDevice, bt->stream));
    FB_HIP(hipStreamSynchronize(bt->stream));   // tEnd is host memory owned by the caller's frame
    return FB_OK;
}

// Forward passes of every member in lockstep: member b runs steps[b] steps
// with seeds[b]; losses[b] receives its loss (may be NULL).
extern "C" fb_status fb_diff_batch_forward(fb_diff_batch *bt, const int32_t *steps, const double *seeds,
                                           int32_t record, double *losses) {
    FB_REQUIRE_D(bt && steps && seeds, FB_ERR_ARG, "fb_diff_batch_forward: bad arguments");
    const int32_t B = (int32_t)bt->m.size();
    hipStream_t st = bt->stream;
    fb_status s;
    int32_t Tmax = 0, Nmax = 0;
    for (int32_t b = 0; b < B; b++) {
        FB_REQUIRE_D(steps[b] >= 0, FB_ERR_ARG, "fb_diff_batch_forward: steps[%d] < 0", b);
        FB_REQUIRE_D(!bt->m[b]->d.nDI && !bt->m[b]->d.dPer, FB_ERR_UNSUPPORTED,
                     "fb_diff_batch_forward: member %d has a drive schedule or per-epoch spike weights; "
                     "run it on its own (fb_diff_forward)", b);
        if ((s = fwd_prepare(bt->m[b], steps[b], seeds[b], record, st))) return s;
        Tmax = std::max(Tmax, steps[b]);
        Nmax = std::max(Nmax, bt->m[b]->N);
    }
    bt->steps.assign(steps, steps + B);
    if ((s = batch_upload(bt, bt->steps)) || (s = batch_ends(bt, bt->steps))) return s;
    const int32_t K = bt->m[0]->d.K;
    const int gN = grid_of(Nmax, 256);
    hipEvent_t e0, e1;
    FB_HIP(hipEventCreate(&e0));
    FB_HIP(hipEventCreate(&e1));
    FB_HIP(hipEventRecord(e0, st));
    bool fused = bt->m[0]->d.D >= 2;
    for (fb_diff *f : bt->m) fused = fused && f->d.nDriven <= 2048;
    for (int32_t t = 0; t < Tmax; t++) {
        if (record && t % K == 0) hipLaunchKernelGGL(kdb_save, dim3(gN, B), dim3(256), 0, st, bt->ds, bt->tEnd, t, t / K);
        if (fused) {
            hipLaunchKernelGGL(kdb_deliver_driven_f, dim3(DELIVER_WG + 1, B), dim3(1024), 0, st, bt->ds, bt->tEnd, t,
                               t);
        } else {
            hipLaunchKernelGGL(kdb_deliver_f, dim3(256, B), dim3(256), 0, st, bt->ds, bt->tEnd, t, t, 0);
            hipLaunchKernelGGL(kdb_driven_f, dim3(1, B), dim3(1024), 0, st, bt->ds, bt->tEnd, t, t);
        }
        hipLaunchKernelGGL(kdb_long_s, dim3(DIFF_LONG_WG, B), dim3(256), 0, st, bt->ds, bt->tEnd, t);
        hipLaunchKernelGGL(kdb_membrane_f, dim3(bt->maxBlk, B), dim3(DIFF_BLOCK), 0, st, bt->ds, bt->tEnd, t, t, 0);
    }
    FB_HIP(hipGetLastError());
    for (int32_t b = 0; b < B; b++)
        if ((s = fwd_finish_launch(bt->m[b], steps[b], st))) return s;
    FB_HIP(hipEventRecord(e1, st));
    FB_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    FB_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    bt->fwdMs = ms;
    for (int32_t b = 0; b < B; b++)
        if ((s = fwd_collect(bt->m[b], steps[b], record, losses ? losses + b : nullptr))) return s;
    return FB_OK;
}

// Backward passes of every member in lockstep; each member's gradients are
// then read with fb_diff_grads / fb_diff_type_grads.
// The members' backward passes in lockstep.  Unlike the single network's
// backward, the replay of a segment does not overlap the sweep of the next:
// these kernels are bound by HBM bandwidth, and running both at once was
// slower (379 ms sequential against 440 ms overlapped, 32 members of the
// full-scale workload).
extern "C" fb_status fb_diff_batch_backward(fb_diff_batch *bt, int32_t truncate) {
    FB_REQUIRE_D(bt && truncate >= 0, FB_ERR_ARG, "fb_diff_batch_backward: bad arguments");
    const int32_t B = (int32_t)bt->m.size();
    hipStream_t st = bt->stream;
    fb_status s;
    int32_t Tmax = 0, Nmax = 0;
    for (int32_t b = 0; b < B; b++) {
        FB_REQUIRE_D(bt->m[b]->steps == bt->steps[b], FB_ERR_STATE,
                     "fb_diff_batch_backward: member %d was not last run by this batch", b);
        if ((s = bwd_prepare(bt->m[b], st))) return s;
        Tmax = std::max(Tmax, bt->steps[b]);
        Nmax = std::max(Nmax, bt->m[b]->N);
    }
    std::vector<int32_t> none((size_t)B, 0);
    if ((s = batch_upload(bt, none))) return s;
    const int32_t K = bt->m[0]->d.K, D = bt->m[0]->d.D;
    const int gN = grid_of(Nmax, 256);
    const int32_t nseg = (Tmax + K - 1) / K;
    // every segment's per-member ends, uploaded once
    std::vector<int32_t> ends((size_t)std::max(1, nseg) * B);
    std::vector<int32_t> segTo((size_t)std::max(1, nseg));
    for (int32_t sgi = 0; sgi < nseg; sgi++) {
        const int32_t from = sgi * K;
        int32_t to = from;
        // member b takes part in this segment up to its own end
        for (int32_t b = 0; b < B; b++) {
            int32_t e = std::min(bt->steps[b], from + K);
            if (e < from) e = from;
            ends[(size_t)sgi * B + b] = e;
            to = std::max(to, e);
        }
        segTo[sgi] = to;
    }
    if (bt->segEndsCap < (int64_t)ends.size()) {
        if (bt->segEnds) FB_HIP(hipFree(bt->segEnds));
        FB_HIP(hipMalloc(&bt->segEnds, sizeof(int32_t) * ends.size()));
        bt->segEndsCap = (int64_t)ends.size();
    }
    FB_HIP(hipMemcpyAsync(bt->segEnds, ends.data(), sizeof(int32_t) * ends.size(), hipMemcpyHostToDevice, st));
    FB_HIP(hipStreamSynchronize(st));   // the host vector goes out of scope with this call
    hipEvent_t e0, e1;
    FB_HIP(hipEventCreate(&e0));
    FB_HIP(hipEventCreate(&e1));
    FB_HIP(hipEventRecord(e0, st));
    for (int32_t sgi = nseg - 1; sgi >= 0; sgi--) {
        const int32_t *te = bt->segEnds + (size_t)sgi * B;
        const int32_t from = sgi * K, to = segTo[sgi];
        const int32_t resFromT = std::max(from, D) - D;
        hipLaunchKernelGGL(kdb_restore, dim3(gN, B), dim3(256), 0, st, bt->ds, te, from, sgi);
        for (int32_t t = from; t < to; t++) {
            hipLaunchKernelGGL(kdb_deliver_forced_r, dim3(257, B), dim3(256), 0, st, bt->ds, te, t, t, resFromT);
            hipLaunchKernelGGL(kdb_long_s, dim3(DIFF_LONG_WG, B), dim3(256), 0, st, bt->ds, te, t);
            hipLaunchKernelGGL(kdb_membrane_r, dim3(bt->maxBlk, B), dim3(DIFF_BLOCK), 0, st, bt->ds, te, t, t,
                               t - from);
        }
        for (int32_t t = to - 1; t >= from; t--) {
            if (truncate > 0 && (t + 1) % truncate == 0)
                hipLaunchKernelGGL(kdb_bclear, dim3(gN, B), dim3(256), 0, st, bt->ds, te, t, 1);
            hipLaunchKernelGGL(kdb_bneuron, dim3(bt->maxBlk, B), dim3(DIFF_BLOCK), 0, st, bt->ds, te, t, t,
                               t - from);
            hipLaunchKernelGGL(kdb_barrival, dim3(256, B), dim3(256), 0, st, bt->ds, te, t, t, resFromT);
        }
        FB_HIP(hipGetLastError());
    }
    for (int32_t b = 0; b < B; b++)
        if ((s = bwd_finish_launch(bt->m[b], st))) return s;
    FB_HIP(hipEventRecord(e1, st));
    FB_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    FB_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    bt->bwdMs = ms;
    for (int32_t b = 0; b < B; b++) {
        bt->m[b]->bwdMs = 0.0;
        if ((s = check_err(bt->m[b], "fb_diff_batch_backward"))) return s;
    }
    return FB_OK;
}

extern "C" fb_status fb_diff_batch_stat(fb_diff_batch *bt, const char *name, double *value) {
    FB_REQUIRE_D(bt && name && value, FB_ERR_ARG, "fb_diff_batch_stat: null argument");
    if (!std::strcmp(name, "forward_ms")) { *value = bt->fwdMs; return FB_OK; }
    if (!std::strcmp(name, "backward_ms")) { *value = bt->bwdMs; return FB_OK; }
    if (!std::strcmp(name, "members")) { *value = (double)bt->m.size(); return FB_OK; }
    set_error("fb_diff_batch_stat: unknown statistic '%s'", name);
    return FB_ERR_ARG;
}

// dL / d gain_T for type gains (outScale[i] = gain of i's type): the type sum
// of gLogGain[i] / outScale[i], in f64 (src/brainmodel.js:63-100 semantics).
extern "C" fb_status fb_diff_type_grads(fb_diff *f, const int32_t *typeOf, int32_t nTypes, double *gType) {
    FB_REQUIRE_D(f && typeOf && gType && nTypes > 0, FB_ERR_ARG, "fb_diff_type_grads: bad arguments");
    std::vector<float> lg((size_t)f->N);
    FB_HIP(hipStreamSynchronize(f->stream));
    FB_HIP(hipMemcpy(lg.data(), f->d.gLogGain, sizeof(float) * f->N, hipMemcpyDeviceToHost));
    for (int32_t T = 0; T < nTypes; T++) gType[T] = 0.0;
    for (int32_t i = 0; i < f->N; i++) {
        FB_REQUIRE_D(typeOf[i] >= 0 && typeOf[i] < nTypes, FB_ERR_ARG, "fb_diff_type_grads: type of %d out of range", i);
        gType[typeOf[i]] += (double)lg[i] / (double)f->in[FB_D_OUTSCALE][i];
    }
    return FB_OK;
}

extern "C" fb_status fb_diff_get_state(fb_diff *f, int32_t which, float *dst) {
    FB_REQUIRE_D(f && dst, FB_ERR_ARG, "fb_diff_get_state: null argument");
    const DiffDev &d = f->d;
    const float *src = nullptr;
    switch (which) {
        case FB_D_V: src = d.v; break;
        case FB_D_GE: src = d.gE; break;
        case FB_D_GI: src = d.gI; break;
        case FB_D_ADAPT: src = d.adapt; break;
        case FB_D_RES: src = d.res; break;
        case FB_D_REFR: src = d.refr; break;
        case FB_D_SPIKECOUNT: src = d.spikeCount; break;
        case FB_D_INSCALE: src = d.inScale; break;
        default: set_error("fb_diff_get_state: unknown array %d", which); return FB_ERR_ARG;
    }
    FB_HIP(hipStreamSynchronize(f->stream));
    FB_HIP(hipMemcpy(dst, src, sizeof(float) * f->N, hipMemcpyDeviceToHost));
    return FB_OK;
}

extern "C" fb_status fb_diff_spikes(fb_diff *f, int64_t *off, int32_t *idx, float *amp, int64_t cap, int64_t *total) {
    FB_REQUIRE_D(f && f->d.off, FB_ERR_STATE, "fb_diff_spikes: no forward pass");
    FB_HIP(hipStreamSynchronize(f->stream));
    std::vector<int64_t> o((size_t)f->steps + 1);
    FB_HIP(hipMemcpy(o.data(), f->d.off, sizeof(int64_t) * o.size(), hipMemcpyDeviceToHost));
    if (off) std::memcpy(off, o.data(), sizeof(int64_t) * o.size());
    const int64_t tot = o.back();
    const int64_t m = std::min(tot, cap);
    if (m > 0 && idx) FB_HIP(hipMemcpy(idx, f->d.idx, sizeof(int32_t) * m, hipMemcpyDeviceToHost));
    if (m > 0 && amp) FB_HIP(hipMemcpy(amp, f->d.amp, sizeof(float) * m, hipMemcpyDeviceToHost));
    if (total) *total = tot;
    return FB_OK;
}

extern "C" fb_status fb_diff_readout(fb_diff *f, float *s, double *rate) {
    FB_REQUIRE_D(f && f->d.readS, FB_ERR_STATE, "fb_diff_readout: no forward pass");
    const size_t m = (size_t)f->steps * f->d.R;
    FB_HIP(hipStreamSynchronize(f->stream));
    if (m && s) FB_HIP(hipMemcpy(s, f->d.readS, sizeof(float) * m, hipMemcpyDeviceToHost));
    if (m && rate) FB_HIP(hipMemcpy(rate, f->d.rate, sizeof(double) * m, hipMemcpyDeviceToHost));
    return FB_OK;
}

extern "C" fb_status fb_diff_size(const fb_diff *f, int32_t *N, int64_t *E, int32_t *steps, int32_t *R) {
    FB_REQUIRE_D(f && N && E && steps && R, FB_ERR_ARG, "fb_diff_size: bad arguments");
    *N = f->N;
    *E = f->E;
    *steps = f->steps;
    *R = f->d.R;
    return FB_OK;
}

extern "C" fb_status fb_diff_rng(fb_diff *f, uint32_t *rng) {
    FB_REQUIRE_D(f && rng, FB_ERR_ARG, "fb_diff_rng: null argument");
    FB_HIP(hipStreamSynchronize(f->stream));
    FB_HIP(hipMemcpy(rng, f->d.rng, sizeof(uint32_t), hipMemcpyDeviceToHost));
    return FB_OK;
}

extern "C" fb_status fb_diff_set_stream(fb_diff *f, void *stream) {
    FB_REQUIRE_D(f, FB_ERR_ARG, "fb_diff_set_stream: null handle");
    FB_HIP(hipStreamSynchronize(f->stream));
    f->stream = stream ? (hipStream_t)stream : f->ownStream;
    return FB_OK;
}

// forward_ms, backward_ms (HIP events around the last calls), edge_events and
// arrivals (forward deliveries since creation), steps, tape_entries
extern "C" fb_status fb_diff_stat(fb_diff *f, const char *name, double *value) {
    FB_REQUIRE_D(f && name && value, FB_ERR_ARG, "fb_diff_stat: null argument");
    if (!std::strcmp(name, "forward_ms")) { *value = f->fwdMs; return FB_OK; }
    if (!std::strcmp(name, "backward_ms")) { *value = f->bwdMs; return FB_OK; }
    if (!std::strcmp(name, "steps")) { *value = f->steps; return FB_OK; }
    if (!std::strcmp(name, "edge_events") || !std::strcmp(name, "arrivals")) {
        unsigned long long c[2];
        FB_HIP(hipStreamSynchronize(f->stream));
        FB_HIP(hipMemcpy(c, f->d.events, sizeof c, hipMemcpyDeviceToHost));
        *value = (double)c[name[0] == 'e' ? 0 : 1];
        return FB_OK;
    }
    if (!std::strcmp(name, "tape_entries")) {
        int64_t tot = 0;
        if (f->d.off) {
            FB_HIP(hipStreamSynchronize(f->stream));
            FB_HIP(hipMemcpy(&tot, f->d.off + f->steps, sizeof(int64_t), hipMemcpyDeviceToHost));
        }
        *value = (double)tot;
        return FB_OK;
    }
    set_error("fb_diff_stat: unknown statistic '%s'", name);
    return FB_ERR_ARG;
}
