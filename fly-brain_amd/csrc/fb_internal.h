// fb_internal.h -- shared definitions of the HIP engine (not part of the ABI).
//
// Layout in HBM, per brain (all SoA, f32 unless noted; N neurons):
//   state      v gE gI refr trace adapt res bias thr drive  (10 x N), spikeCount u32[N]
//   delay line ring i32[nslots x 2N] (one spike list per slot; 2N because with
//              tRef <= 0 a forced neuron may also cross threshold in its step),
//              ringCount i32[nslots]
//   delivery   cnt i32[2N]   same-step arrivals per (sign class, target)
//              add f32[2N]   the first arrival's addend f32(W*s)
//              key u64[2N]   its order key (arrival position << 32 | position in row)
//              runKey u64[2E], runAdd f32[2E]  further arrivals: class c of target q
//                            owns entries c*E + inOff[q] .. c*E + inOff[q+1] - 1 (each
//                            in-edge arrives at most once per step)
//   background bgCnt i32[N]  background hits this step (all add the same bgAmp)
//   compaction stage i32[N], blkCnt i32[nblocks]
// Per graph: indptr u32[N+1], edges {u32 target, f32 W}[E] interleaved so one
// 8-byte load feeds one edge, sign f32[N], counts i32[E] (optional), and the
// in-degree prefix inOff u32[N+1].
#pragma once
#include <hip/hip_runtime.h>
#include <mutex>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/flybrain.h"

namespace fb {

void set_error(const char *fmt, ...);

#define FB_HIP(call)                                                                 \
    do {                                                                             \
        hipError_t _e = (call);                                                      \
        if (_e != hipSuccess) {                                                      \
            ::fb::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(_e),   \
                            __FILE__, __LINE__);                                     \
            return FB_ERR_DEVICE;                                                    \
        }                                                                            \
    } while (0)

// f32 header constants, each formed in f64 and rounded once (src/lifwasm.js:58-67)
struct Consts {
    float dt, vRest, vThresh, vReset, tRef, adaptInc, depU;
    float dE, dTr, dA, kRec, kM, dtS, eExc, eInh, cE, cI;
    float bgEvents, bgAmp, preInh;
    int32_t coba;
};

// device-side scalars of one brain
struct Scalars {
    uint32_t rng;
    float bgAcc;
    int32_t nDrivenFired;   // forced spikes emitted this step
    int32_t err;            // sticky error bits (FB_DEVERR_*)
    int32_t nLong;          // (class, target) pairs with more than LONG_RUN arrivals this step
    int32_t pad;
    unsigned long long edgeEvents;
    unsigned long long synEvents;
    unsigned long long spikes;
};
enum { FB_DEVERR_RUN_OVERFLOW = 1 };

struct DevGraph {
    int32_t N;
    int64_t E;
    const uint32_t *indptr;
    const uint2 *edges;      // {target, bits of W}
    const float *sign;
    const int32_t *counts;
    const uint32_t *rowSyn;  // synapse count summed over each row (0 without counts)
    const uint32_t *inOff;   // [N+1] in-degree prefix: target q owns run entries inOff[q] .. inOff[q+1]-1
    int64_t maxRow;          // longest row
    // host memory, never read on the device: per 64-neuron chunk, the work it
    // brings the lookahead engine (non-zero in-edges + non-zero out-edges +
    // 16 per neuron), which its regular blocks balance
    const uint64_t *chunkCost;
};

struct DevBrain {
    int32_t N, nslots, nDriven, nblocks;
    int32_t skipZeroW;         // deliveries may skip zero-weight edges (exact: see fb_brain::gENonneg)
    int64_t ringStride;        // entries per ring slot (2N)
    Consts k;
    DevGraph g;
    float *v, *gE, *gI, *refr, *trace, *adapt, *res, *bias, *thr, *drive;
    uint32_t *spikeCount;
    int32_t *ring, *ringCount, *driven;
    int32_t *cnt;              // [2N] arrivals this step per (sign class, target)
    float *add;                // [2N] the first arrival's addend and order key
    uint64_t *key;
    uint64_t *runKey;          // [2E] later arrivals: class c, target q at c*E + inOff[q] + (slot-1)
    float *runAdd;
    int32_t *longList;         // [2N] class*N + target of the step's long runs (k_long sums them)
    int32_t *bgCnt;
    int32_t *stage, *blkCnt;
    Scalars *sc;
    const uint32_t *jumpTab;   // 32 powers x 4 bytes x 256 entries of xorshift32 jumps
    const uint32_t *leap;      // LEAP_MATS leapfrog matrices (see leap_fill)
    // the closed pool loop (fb_set_pool_feedback); poolOf null when off
    const uint8_t *poolOf;     // [N] pool of each neuron, or >= loopPools for none
    const int32_t *loopTargets;
    uint32_t *loopHist;        // [loopRing] pool totals by step
    uint16_t *loopRec;         // [loopRecCap][loopPools] per-step pool counts
    int32_t *loopState;        // [0] steps since the loop was set, [1] records written
    int32_t loopPools, loopNTargets, loopDelay, loopRing, loopRecCap;
    float loopBase, loopGain;
};

// ------------------------------------------------------------------ device helpers

__device__ __forceinline__ uint32_t xs32(uint32_t x) {
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return x;
}

// Applies M^(2^b) (M = one xorshift32 step as a GF(2) matrix) by 4 byte-table lookups.
__device__ __forceinline__ uint32_t jump_pow2(const uint32_t *tab, int b, uint32_t x) {
    const uint32_t *t = tab + (size_t)b * 1024;
    return t[x & 255u] ^ t[256 + ((x >> 8) & 255u)] ^ t[512 + ((x >> 16) & 255u)] ^ t[768 + (x >> 24)];
}

// State after n further xorshift32 steps.
__device__ __forceinline__ uint32_t jump(const uint32_t *tab, uint32_t x, uint32_t n) {
    while (n) {
        int b = __ffs(n) - 1;
        x = jump_pow2(tab, b, x);
        n &= n - 1;
    }
    return x;
}

// ------------------------------------------------------------------ parallel stream generation
//
// lif.c draws from one serial xorshift32 stream.  xorshift32 is linear over
// GF(2): one step is a 32x32 bit matrix M, so the state p steps ahead is
// M^p x.  A workgroup of 1024 threads produces 2048 consecutive states per
// round: thread t computes positions t+1 and t+1025 as
//   M^(t%32 + 1) (M^(32 h) x),   h = t/32 and t/32 + 32,
// from 96 matrices: M^1..M^32 (fine) then M^0, M^32, ..., M^2016 (coarse).
// A matrix is kept as 8 nibble tables (16 words each: the XOR of the columns
// selected by a 4-bit slice of x), so a product is 8 LDS reads.  In LDS the
// matrices sit LEAP_STRIDE = 129 words apart, so lanes that read the same
// (nibble, value) entry of 32 different matrices hit 32 different banks.
constexpr int LEAP_MATS = 96;
// The global table holds one more matrix after them, M^2048, for jumps past
// the first 2048 stream positions (it is not staged in LDS).
constexpr int LEAP_FAR = 96;
constexpr int LEAP_TABLE_MATS = 97;
constexpr int LEAP_WORDS = 128;           // 8 nibbles x 16 values
constexpr int LEAP_ROUND = 2048;
constexpr int LEAP_STRIDE = 129;

// Grid-wide barrier for the cooperative (all workgroups resident) kernels.
// The runtime's grid sync costs ~25 us on MI355X when ~250 workgroups arrive
// together; this one costs ~7 us there and ~2 us when one workgroup is late.
// Workgroups arrive on their group's counter (GBAR_GROUP a group, counters 256
// bytes apart), the last of a group on the top counter ctr[0], and every
// workgroup polls the top counter.  The counters only grow, so a barrier needs
// no reset; the host zeroes them before each launch.  Memory: every wave waits
// for its own accesses, then one agent-scope release fence and, after the
// wait, one agent-scope acquire fence per workgroup.  The wait is bounded: after
// GBAR_TIMEOUT wall-clock ticks it sets bit 8 in *flags and returns false, and
// the caller leaves the kernel.
// Kernels with grid barriers need every workgroup resident at once.  Their
// grids hold at most one workgroup per CU at the kernel's occupancy, and the
// library runs one of them at a time per process: this lock is held from a
// launch until the host has seen the kernel end (the lookahead engine and the
// ensemble lookahead engine).  Kernels of other streams are finite, so their
// CUs free up; and a barrier wait is bounded, so a grid that cannot become
// resident (another process's grid-barrier kernel) ends in an error.
std::mutex &grid_launch_mutex();

constexpr int GBAR_GROUP = 16;
constexpr int GBAR_MAX_BLOCKS = 1024;
constexpr int GBAR_WORDS = 64 * (GBAR_MAX_BLOCKS / GBAR_GROUP + 1);
constexpr uint64_t GBAR_TIMEOUT = 1000000000ull;   // 10 s of the 100 MHz wall clock
constexpr int32_t GBAR_TIMED_OUT = 8;

__device__ __forceinline__ bool grid_barrier(int32_t *ctr, int32_t *flags) {
    __shared__ int32_t gbOk;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int32_t nblk = (int32_t)gridDim.x;
        const int32_t g = (int32_t)blockIdx.x / GBAR_GROUP, ng = (nblk + GBAR_GROUP - 1) / GBAR_GROUP;
        const int32_t sz = min(GBAR_GROUP, nblk - g * GBAR_GROUP);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const int32_t arr = atomicAdd(&ctr[64 * (g + 1)], 1);
        if (arr % sz == sz - 1) atomicAdd(&ctr[0], 1);
        const int32_t target = (arr / sz + 1) * ng;
        const uint64_t t0 = wall_clock64();
        int32_t ok = 1;
        while (__hip_atomic_load(&ctr[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(1);
            if (wall_clock64() - t0 > GBAR_TIMEOUT) {
                atomicOr(flags, GBAR_TIMED_OUT);
                ok = 0;
                break;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        gbOk = ok;
    }
    __syncthreads();
    return gbOk != 0;
}

// Lookahead-kernel clocks reported per launch: the NPH phases of the driven
// block, of the first regular block and of the first spare block; the parts of
// the driven step (6); the shared push of the driven rows (mean load / edges /
// publish over blocks, then the slowest block's total); the launch's span
// (7 clocks: ~first block start, last block start, last block end, the last
// block out's final write, the last end of a driven, a regular and a spare
// block; the first is stored complemented so that every slot is an atomic
// maximum over a zeroed word).
constexpr int NPH = 8;   // gather, sim, push, lists, barrier, prologue, share, the wait for the pushes before the share
constexpr int PROF_SUB = 3 * NPH;
constexpr int PROF_SHARE = PROF_SUB + 6;
constexpr int PROF_SPAN = PROF_SHARE + 4;
constexpr int PERSIST_PROF = PROF_SPAN + 7;
constexpr int LEAP_LDS_WORDS = LEAP_MATS * LEAP_STRIDE;

__device__ __forceinline__ uint32_t matvec(const uint32_t *tab, uint32_t x) {
    uint32_t r = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) r ^= tab[k * 16 + ((x >> (4 * k)) & 15u)];
    return r;
}

// global layout (stride 128) -> LDS layout (stride 129), the first MATS
// matrices; every thread of the block calls it
template <int T, int MATS = LEAP_MATS>
__device__ __forceinline__ void leap_load(uint32_t *leapS, const uint32_t *leap) {
    // every load issued before the first store: one memory round trip, not
    // one per word a thread copies
    constexpr int WORDS = MATS * LEAP_WORDS, PER = (WORDS + T - 1) / T;
    uint32_t v[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int q = (int)threadIdx.x + k * T;
        v[k] = q < WORDS ? leap[q] : 0u;
    }
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int q = (int)threadIdx.x + k * T;
        if (q < WORDS) leapS[(q >> 7) * LEAP_STRIDE + (q & 127)] = v[k];
    }
}

// M^J x for 1 <= J from LDS tables holding all LEAP_TABLE_MATS matrices:
// with J - 1 = 2048 a + 32 h + f, M^J = M^(f+1) M^(32 h) (M^2048)^a.
__device__ __forceinline__ uint32_t leap_jump(const uint32_t *leapS, uint32_t J, uint32_t x) {
    const uint32_t o = J - 1u;
    for (uint32_t a = o >> 11; a; a--) x = matvec(leapS + LEAP_FAR * LEAP_STRIDE, x);
    return matvec(leapS + (o & 31u) * LEAP_STRIDE, matvec(leapS + (32u + ((o >> 5) & 63u)) * LEAP_STRIDE, x));
}

// Exclusive scan over a T-thread workgroup of per-thread counts given as K
// flag bits (bit k of `bits`): ballots and popcounts inside a wave, wave
// totals through LDS.  sh needs T/64 entries; one barrier.  Every thread must
// call it, and sh must not be rewritten before every thread has returned.
template <int T, int K>
__device__ __forceinline__ int32_t wg_scan_flags(uint32_t bits, int32_t *sh, int32_t *total) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t lt = (1ull << lane) - 1ull;
    int32_t inWave = 0, waveTotal = 0;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const uint64_t mk = __ballot((bits >> k) & 1u);
        inWave += __popcll(mk & lt);
        waveTotal += __popcll(mk);
    }
    if (lane == 0) sh[wave] = waveTotal;
    __syncthreads();
    int32_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < T / 64; w++) {
        const int32_t c = sh[w];
        before += w < wave ? c : 0;
        all += c;
    }
    *total = all;
    return before + inWave;
}

// Exclusive scan of one value per thread over a T-thread workgroup (T a
// multiple of 64, at most 4096): wave shuffles, then one wave over the wave
// totals.  sh needs 64 + T/64 entries.  Every thread must call it.
template <int T>
__device__ inline int32_t wg_exclusive_scan(int32_t v, int32_t *sh, int32_t *total) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) sh[wave] = x;
    __syncthreads();
    if (tid < 64) {
        int32_t w = tid < T / 64 ? sh[tid] : 0;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int32_t y = __shfl_up(w, off, 64);
            if (lane >= off) w += y;
        }
        if (tid < T / 64) sh[64 + tid] = w;
    }
    __syncthreads();
    const int32_t waveBase = wave ? sh[64 + wave - 1] : 0;
    *total = sh[64 + T / 64 - 1];
    __syncthreads();
    return waveBase + x - v;
}

// Fills out[0, min(n - r0, 2048)) with the stream states at positions
// r0+1 .. r0+2048 after `base`, where base is the state at position r0.
// leapS: the 96 matrices in LDS; ybase: 64 words of LDS.  1024 threads.
__device__ inline void leap_round(const uint32_t *leapS, uint32_t *ybase, uint32_t *out, int32_t count,
                                  uint32_t base) {
    const int tid = threadIdx.x;
    if (tid < 64) ybase[tid] = matvec(leapS + (32 + tid) * LEAP_STRIDE, base);
    __syncthreads();
    const uint32_t *fine = leapS + (tid & 31) * LEAP_STRIDE;
    if (tid < count) out[tid] = matvec(fine, ybase[tid >> 5]);
    if (tid + 1024 < count) out[tid + 1024] = matvec(fine, ybase[(tid >> 5) + 32]);
    __syncthreads();
}

// Membrane update of one neuron (src/wasm/lif.c:57-71): the reference's C
// association, one f32 rounding per operation (the library is built with
// -ffp-contract=off so no multiply-add is fused).
struct NeuronState {
    float v, gE, gI, refr, trace, adapt, res;
};

__device__ __forceinline__ void membrane(const Consts &k, NeuronState &s, float bias) {
    const float vi = s.v, r = s.refr;
    float drv;
    if (k.coba) {
        const float exc = (s.gE * (k.eExc - vi)) * k.cE;
        const float inh = (s.gI * (vi - k.eInh)) * k.cI;
        drv = (((k.vRest - vi) + exc) + inh) + bias;
    } else {
        drv = (((k.vRest - vi) + s.gE) + s.gI) + bias;
    }
    const float dv = drv * k.kM;
    s.v = r > 0 ? k.vReset : vi + dv;
    s.refr = r > 0 ? r - k.dt : r;
    s.gE = s.gE * k.dE;
    s.gI = s.gI * k.dE;
    s.trace = s.trace * k.dTr;
    s.adapt = s.adapt * k.dA;
    s.res = s.res + (1.f - s.res) * k.kRec;
}

// lif.c:75 -- threshold test against the decayed adapt and the per-neuron offset
__device__ __forceinline__ bool crosses(const Consts &k, const NeuronState &s, float thr) {
    return s.v >= (k.vThresh + s.adapt) + thr && s.refr <= 0;
}

__device__ __forceinline__ void spike_effects(const Consts &k, NeuronState &s, float refrValue) {
    s.v = k.vReset;
    s.refr = refrValue;
    s.trace = 1.f;
    s.adapt = s.adapt + k.adaptInc;
}

// Sum of the same-step arrivals on one (sign class, target) in arrival order:
// the first arrival is (key0, add0), the other n-1 sit in rk/ra[0, n-1) in any
// order.  Keys are distinct.  Up to M arrivals: a bitonic network over
// registers (missing entries act as +infinity), then the ordered sum.
template <int M>
__device__ __forceinline__ float ordered_sum_net(float g, int32_t n, uint64_t key0, float add0, const uint64_t *rk,
                                                 const float *ra) {
    uint64_t k[M];
    float a[M];
    k[0] = key0;
    a[0] = add0;
#pragma unroll
    for (int q = 1; q < M; q++) {
        const bool has = q < n;
        k[q] = has ? rk[q - 1] : ~0ull;
        a[q] = has ? ra[q - 1] : 0.f;
    }
#pragma unroll
    for (int size = 2; size <= M; size <<= 1) {
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
#pragma unroll
            for (int i = 0; i < M; i++) {
                const int j = i ^ stride;
                if (j > i) {
                    const bool up = (i & size) == 0;
                    if ((k[i] > k[j]) == up) {
                        const uint64_t tk = k[i]; k[i] = k[j]; k[j] = tk;
                        const float ta = a[i]; a[i] = a[j]; a[j] = ta;
                    }
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < M; q++)
        if (q < n) g = g + a[q];
    return g;
}

// Runs of up to LONG_RUN arrivals are summed by the neuron's own thread;
// longer ones by a workgroup each (k_long) before the membrane pass.
constexpr int LONG_RUN = 16;

__device__ inline float ordered_sum(float g, int32_t n, uint64_t key0, float add0, const uint64_t *rk,
                                    const float *ra) {
    if (n <= 4) return ordered_sum_net<4>(g, n, key0, add0, rk, ra);
    if (n <= 8) return ordered_sum_net<8>(g, n, key0, add0, rk, ra);
    return ordered_sum_net<LONG_RUN>(g, n, key0, add0, rk, ra);
}

}  // namespace fb
