// fb_persistent.hip -- the lookahead engine: one cooperative launch advances
// many LIF steps, with the delay line as a lookahead window.
//
// A spike emitted at step t is delivered at step t + D (D = nslots - 1, 18 at
// dt 0.1 ms).  So at the start of an epoch [t0, t0+L) with L <= D, every
// arrival of the epoch is already in the ring, and each neuron evolves
// independently of every other neuron until the epoch ends -- except for the
// forced spikes, which draw from one serial xorshift32 stream in driven-list
// order.  The launch therefore splits the grid into:
//
//   block 0         the driven block: owns every driven neuron (state in
//                   registers, list order across threads), draws the stream
//                   per step with a workgroup scan and a two-level leapfrog:
//                   draw t+1 of a step is M^(t%32+1) M^(32*(t/32)) x, where M is
//                   one xorshift32 step as a 32x32 matrix over GF(2) and the 64
//                   matrices sit in LDS.
//   blocks 1..G-1   regular blocks: block b owns a contiguous range of up to
//                   2048 neurons (2 per thread, state in registers).
//
// Per epoch, three grid barriers:
//   E1  each regular block binary-searches every arriving row for the slice
//       of targets in its range (rows are sorted by target) and keeps those
//       edge events in LDS, sorted by (step, sign class, target, arrival
//       position, position in row) -- lif.c's addition order per target.
//       Events for driven targets are forwarded to the driven block.
//   E2  every block advances its neurons L steps from registers: arrivals
//       (one add, or the sorted run for several), forced spikes, membrane,
//       threshold.  Threshold spikes go to a per-step bitmask (wave ballots),
//       forced spikes to per-step flags in list order.
//   E3  block m materialises step m's spike list into its ring slot in
//       lif.c's order (forced, then ascending) by a workgroup scan.
//
// Preconditions (checked on the host; otherwise the caller uses the per-step
// engines, which accept every parameter set): no background events, no
// presynaptic inhibition, at most 2048 driven neurons, at most 2048 neurons
// per block.  With synaptic depression live the epoch is also capped at the
// refractory period, so a neuron arrives at most once per epoch and its
// depression bookkeeping stays local.  If an
// epoch's events overflow the LDS, the launch stops at the epoch boundary and
// reports how many steps it completed; the host runs that epoch with the
// per-step engine and relaunches.
#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <map>

#include "fb_internal.h"

namespace cg = cooperative_groups;

namespace fb {

constexpr int PT = 1024;            // threads per block
constexpr int NPT = 2;              // neurons per thread in a regular block
constexpr int NB_MAX = PT * NPT;    // neurons per regular block
constexpr int D_MAX = 32;           // epoch length bound
constexpr int EV_CAP = 4096;        // LDS events per block per epoch
constexpr int DPT = 2;              // driven neurons per thread in block 0 (wg_scan_flags2 takes two)
constexpr int DRV_CAP = PT * DPT;
constexpr int WORDS_PER_STEP = NB_MAX / 32;

// event key: step(5) | class(1) | local target(12) | arrival position(20) | position in row(26)
__device__ __forceinline__ uint64_t ev_key(uint32_t m, uint32_t c, uint32_t ql, uint32_t k, uint32_t jr) {
    return ((uint64_t)m << 59) | ((uint64_t)c << 58) | ((uint64_t)ql << 46) | ((uint64_t)k << 26) | (uint64_t)jr;
}
__device__ __forceinline__ uint32_t ev_group(uint64_t key) { return (uint32_t)(key >> 46); }
__device__ __forceinline__ uint32_t group_of(uint32_t m, uint32_t c, uint32_t ql) { return (m << 13) | (c << 12) | ql; }

struct PersistArgs {
    DevBrain d;
    int32_t head0;       // ring head at launch
    int32_t nsteps;
    int32_t L;           // epoch length (<= D, <= refractory length)
    int32_t nb;          // neurons per regular block
    int32_t nwords;      // 64-bit words of one step's spike bitmask
    uint64_t *spk;       // [D_MAX][nwords]
    uint8_t *dflag;      // [D_MAX][nDriven]
    const int32_t *drvPos;   // [N] position in the driven list, -1 if not driven
    uint64_t *fwdKey;    // events forwarded to the driven block
    float *fwdAdd;
    int32_t fwdCap;
    int32_t *ctl;        // [0] steps done, [1] abort flag, [2] forwarded count
    uint64_t *prof;      // per-phase clock ticks of blocks 0 and 1 (12), driven-block step parts (6)
};

struct Shared {
    uint64_t evKey[EV_CAP];
    float evAdd[EV_CAP];
    uint32_t hasEv[D_MAX * WORDS_PER_STEP];
    uint32_t arr[D_MAX * WORDS_PER_STEP];
    uint32_t draws[DRV_CAP];
    uint32_t leap[LEAP_LDS_WORDS];
    uint32_t ybase[64];
    int32_t scan[64 + PT / 64];
    uint64_t ph[7];      // phase clocks of thread 0: six accumulators and the last stamp
    uint64_t sub[7];     // driven-block step clocks: five accumulators and the last stamp
    int32_t nev;
    int32_t pref[D_MAX + 1];
    uint32_t rngBase;
};

// ------------------------------------------------------------------ small helpers

__device__ __forceinline__ uint32_t lower_bound_target(const uint2 *edges, uint32_t a, uint32_t e, uint32_t v) {
    while (a < e) {
        const uint32_t mid = a + ((e - a) >> 1);
        if (edges[mid].x < v) a = mid + 1; else e = mid;
    }
    return a;
}

// LDS bitonic sort of evKey/evAdd[0, n), padded to a power of two with ~0 keys
__device__ inline void lds_sort(Shared &s, int32_t n) {
    int32_t P = 1;
    while (P < n) P <<= 1;
    for (int32_t i = n + threadIdx.x; i < P; i += PT) { s.evKey[i] = ~0ull; s.evAdd[i] = 0.f; }
    __syncthreads();
    for (int32_t k = 2; k <= P; k <<= 1) {
        for (int32_t j = k >> 1; j > 0; j >>= 1) {
            for (int32_t i = threadIdx.x; i < P; i += PT) {
                const int32_t l = i ^ j;
                if (l > i) {
                    const bool up = (i & k) == 0;
                    const uint64_t ki = s.evKey[i], kl = s.evKey[l];
                    if ((ki > kl) == up) {
                        s.evKey[i] = kl; s.evKey[l] = ki;
                        const float t = s.evAdd[i]; s.evAdd[i] = s.evAdd[l]; s.evAdd[l] = t;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// first index in evKey[0, n) whose group is >= g
__device__ __forceinline__ int32_t find_group(const Shared &s, int32_t n, uint32_t g) {
    int32_t a = 0, e = n;
    while (a < e) {
        const int32_t mid = (a + e) >> 1;
        if (ev_group(s.evKey[mid]) < g) a = mid + 1; else e = mid;
    }
    return a;
}

__device__ __forceinline__ float apply_group(const Shared &s, int32_t n, uint32_t g, float acc) {
    for (int32_t q = find_group(s, n, g); q < n && ev_group(s.evKey[q]) == g; q++) acc = acc + s.evAdd[q];
    return acc;
}

// res at the start of epoch step m from its value at the epoch start: m recoveries
__device__ __forceinline__ float res_after(const Consts &k, float r, int32_t m) {
    for (int32_t j = 0; j < m; j++) r = r + (1.f - r) * k.kRec;
    return r;
}

__device__ __forceinline__ int32_t ring_slot(const PersistArgs &a, int32_t done, int32_t m) {
    return (a.head0 + done + m) % a.d.nslots;      // head of epoch step m
}

// ------------------------------------------------------------------ E1: arrivals -> events

__device__ void e1_regular(const PersistArgs &a, Shared &s, int32_t done, int32_t L, int32_t lo, int32_t hi) {
    const DevBrain &d = a.d;
    const int tid = threadIdx.x;
    const int32_t words = (a.nb + 31) >> 5;
    for (int32_t w = tid; w < L * WORDS_PER_STEP; w += PT) { s.hasEv[w] = 0u; s.arr[w] = 0u; }
    if (tid == 0) {
        s.nev = 0;
        int32_t acc = 0;
        for (int32_t m = 0; m < L; m++) { s.pref[m] = acc; acc += d.ringCount[ring_slot(a, done, m)]; }
        s.pref[L] = acc;
    }
    __syncthreads();
    (void)words;
    const int32_t Atot = s.pref[L];
    for (int32_t g = tid; g < Atot; g += PT) {
        int32_t m = 0;
        while (s.pref[m + 1] <= g) m++;
        const int32_t k = g - s.pref[m];
        const int32_t pre = d.ring[(int64_t)ring_slot(a, done, m) * d.ringStride + k];
        const float r = res_after(d.k, d.res[pre], m);
        const float sv = d.g.sign[pre] * r;
        if (pre >= lo && pre < hi && a.drvPos[pre] < 0 && sv != 0.f) {
            const int32_t ql = pre - lo;
            atomicOr(&s.arr[m * WORDS_PER_STEP + (ql >> 5)], 1u << (ql & 31));
        }
        if (sv == 0.f) continue;
        const uint32_t ra = d.g.indptr[pre], re = d.g.indptr[pre + 1];
        const uint32_t j0 = lower_bound_target(d.g.edges, ra, re, (uint32_t)lo);
        const uint32_t j1 = lower_bound_target(d.g.edges, j0, re, (uint32_t)hi);
        if (j1 == j0) continue;
        const uint32_t c = sv > 0 ? 0u : 1u;
        const int32_t base = atomicAdd(&s.nev, (int32_t)(j1 - j0));
        for (uint32_t j = j0; j < j1; j++) {
            const uint2 ed = d.g.edges[j];
            const float addend = __uint_as_float(ed.y) * sv;
            const int32_t q = (int32_t)ed.x;
            const int32_t idx = base + (int32_t)(j - j0);
            const int32_t dp = a.drvPos[q];
            if (dp >= 0) {
                const int32_t p = atomicAdd(&a.ctl[2], 1);
                if (p < a.fwdCap) {
                    a.fwdKey[p] = ev_key((uint32_t)m, c, (uint32_t)dp, (uint32_t)k, j - ra);
                    a.fwdAdd[p] = addend;
                } else {
                    atomicOr(&a.ctl[1], 1);
                }
                if (idx < EV_CAP) { s.evKey[idx] = ~0ull; s.evAdd[idx] = 0.f; }
            } else if (idx < EV_CAP) {
                const uint32_t ql = (uint32_t)(q - lo);
                s.evKey[idx] = ev_key((uint32_t)m, c, ql, (uint32_t)k, j - ra);
                s.evAdd[idx] = addend;
                atomicOr(&s.hasEv[m * WORDS_PER_STEP + (ql >> 5)], 1u << (ql & 31));
            }
        }
    }
    __syncthreads();
    if (tid == 0 && s.nev > EV_CAP) atomicOr(&a.ctl[1], 2);
}

// the driven block: arrival marks for its own neurons and the work counters
__device__ void e1_driven(const PersistArgs &a, Shared &s, int32_t done, int32_t L, unsigned long long &ev,
                          unsigned long long &syn) {
    const DevBrain &d = a.d;
    const int tid = threadIdx.x;
    for (int32_t w = tid; w < L * WORDS_PER_STEP; w += PT) s.arr[w] = 0u;
    if (tid == 0) {
        int32_t acc = 0;
        for (int32_t m = 0; m < L; m++) { s.pref[m] = acc; acc += d.ringCount[ring_slot(a, done, m)]; }
        s.pref[L] = acc;
    }
    __syncthreads();
    const int32_t Atot = s.pref[L];
    ev = 0;
    syn = 0;
    for (int32_t g = tid; g < Atot; g += PT) {
        int32_t m = 0;
        while (s.pref[m + 1] <= g) m++;
        const int32_t k = g - s.pref[m];
        const int32_t pre = d.ring[(int64_t)ring_slot(a, done, m) * d.ringStride + k];
        const float sv = d.g.sign[pre] * res_after(d.k, d.res[pre], m);
        if (sv == 0.f) continue;
        ev += d.g.indptr[pre + 1] - d.g.indptr[pre];
        syn += d.g.rowSyn[pre];
        const int32_t dp = a.drvPos[pre];
        if (dp >= 0) atomicOr(&s.arr[m * WORDS_PER_STEP + (dp >> 5)], 1u << (dp & 31));
    }
    __syncthreads();
}

// ------------------------------------------------------------------ E2: L steps from registers

struct Cell {
    NeuronState st;
    float bias, thr, sign;
    uint32_t count;   // spike count increments this launch
    int32_t i;        // neuron index, -1 if the slot is empty
};

__device__ __forceinline__ void load_cell(const DevBrain &d, Cell &c, int32_t i) {
    c.i = i;
    c.count = 0;
    if (i < 0) return;
    c.st.v = d.v[i]; c.st.gE = d.gE[i]; c.st.gI = d.gI[i]; c.st.refr = d.refr[i];
    c.st.trace = d.trace[i]; c.st.adapt = d.adapt[i]; c.st.res = d.res[i];
    c.bias = d.bias[i]; c.thr = d.thr[i]; c.sign = d.g.sign[i];
}

__device__ __forceinline__ void store_cell(const DevBrain &d, Cell &c) {
    if (c.i < 0) return;
    const int32_t i = c.i;
    d.v[i] = c.st.v; d.gE[i] = c.st.gE; d.gI[i] = c.st.gI; d.refr[i] = c.st.refr;
    d.trace[i] = c.st.trace; d.adapt[i] = c.st.adapt; d.res[i] = c.st.res;
    d.spikeCount[i] += c.count;
    c.count = 0;
}

// lif.c phase 1 for one neuron: depression if it arrives now, then this step's inputs
__device__ __forceinline__ void cell_inputs(const Consts &k, Cell &c, const Shared &s, int32_t nev, int32_t m,
                                            uint32_t ql) {
    if ((s.arr[m * WORDS_PER_STEP + (ql >> 5)] >> (ql & 31)) & 1u) {
        if (c.sign * c.st.res != 0.f) c.st.res = c.st.res - k.depU * c.st.res;
    }
    if ((s.hasEv[m * WORDS_PER_STEP + (ql >> 5)] >> (ql & 31)) & 1u) {
        c.st.gE = apply_group(s, nev, group_of((uint32_t)m, 0u, ql), c.st.gE);
        c.st.gI = apply_group(s, nev, group_of((uint32_t)m, 1u, ql), c.st.gI);
    }
}

__device__ void e2_regular(const PersistArgs &a, Shared &s, Cell (&cells)[NPT], int32_t done, int32_t L,
                           int32_t lo) {
    const DevBrain &d = a.d;
    const int32_t nev = min(s.nev, EV_CAP);
    if (nev > 1) lds_sort(s, nev);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    for (int32_t m = 0; m < L; m++) {
#pragma unroll
        for (int u = 0; u < NPT; u++) {
            Cell &c = cells[u];
            bool spk = false;
            if (c.i >= 0) {
                cell_inputs(d.k, c, s, nev, m, (uint32_t)(c.i - lo));
                membrane(d.k, c.st, c.bias);
                if (crosses(d.k, c.st, c.thr)) {
                    spike_effects(d.k, c.st, d.k.tRef);
                    c.count++;
                    spk = true;
                }
            }
            const uint64_t mask = __ballot(spk);
            // the wave's 64 neurons are consecutive and 64-aligned
            if (lane == 0 && mask) {
                const int32_t first = lo + (int32_t)(threadIdx.x & ~63) + u * PT;
                atomicOr((unsigned long long *)&a.spk[(int64_t)m * a.nwords + (first >> 6)], (unsigned long long)mask);
            }
        }
    }
    (void)done;
}

__device__ void e2_driven(const PersistArgs &a, Shared &s, Cell (&cells)[DPT], float (&pFire)[DPT], int32_t L,
                          uint32_t &rng) {
    const DevBrain &d = a.d;
    const int tid = threadIdx.x;
    const int32_t nD = d.nDriven;
    const int32_t P = (nD + PT - 1) / PT;
    // forwarded events
    const int32_t nfw = min(a.ctl[2], a.fwdCap);
    for (int32_t w = tid; w < L * WORDS_PER_STEP; w += PT) s.hasEv[w] = 0u;
    __syncthreads();
    const int32_t nev = min(nfw, EV_CAP);
    for (int32_t q = tid; q < nev; q += PT) {
        const uint64_t key = a.fwdKey[q];
        s.evKey[q] = key;
        s.evAdd[q] = a.fwdAdd[q];
        const uint32_t g = ev_group(key);
        const uint32_t m = g >> 13, ql = g & 4095u;
        atomicOr(&s.hasEv[m * WORDS_PER_STEP + (ql >> 5)], 1u << (ql & 31));
    }
    __syncthreads();
    if (nev > 1) lds_sort(s, nev);
    __syncthreads();
    const float forcedRefr = d.k.tRef + d.k.dt;
    const bool timed = tid == 0;
#define FB_SUB(k)                                      \
    if (timed) {                                       \
        const uint64_t now = wall_clock64();           \
        s.sub[k] += now - s.sub[6];                    \
        s.sub[6] = now;                                \
    }
    if (timed) s.sub[6] = wall_clock64();
    for (int32_t m = 0; m < L; m++) {
        // lif.c phase 1: arrivals for the driven neurons
        int32_t draws = 0;
        uint32_t nonRefr = 0;
#pragma unroll
        for (int u = 0; u < DPT; u++) {
            Cell &c = cells[u];
            if (u < P && c.i >= 0) {
                cell_inputs(d.k, c, s, nev, m, (uint32_t)(tid * P + u));
                if (!(c.st.refr > 0)) { nonRefr |= 1u << u; draws++; }
            }
        }
        (void)draws;
        FB_SUB(0)
        // lif.c phase 2: one draw per non-refractory driven neuron, in list order
        int32_t n;
        const int32_t off = wg_scan_flags2<PT>(nonRefr & 1u, (nonRefr >> 1) & 1u, s.scan + (m & 1) * 16, &n);
        FB_SUB(1)
        // n <= 2048 driven neurons: one leapfrog round covers the step
        if (n > 0) {
            leap_round(s.leap, s.ybase, s.draws, n, rng);
            rng = s.draws[n - 1];
        }
        FB_SUB(2)
        int32_t pos = off;
#pragma unroll
        for (int u = 0; u < DPT; u++) {
            Cell &c = cells[u];
            uint8_t f = 0;
            if ((nonRefr >> u) & 1u) {
                const uint32_t x = s.draws[pos++];
                const float uu = (float)(x >> 8) * (1.0f / 16777216.0f);
                if (uu < pFire[u]) {
                    c.st.v = d.k.vReset;
                    c.st.refr = forcedRefr;
                    c.st.trace = 1.f;
                    c.st.adapt = c.st.adapt + d.k.adaptInc;
                    c.count++;
                    f = 1;
                }
            }
            if (u < P && tid * P + u < nD) a.dflag[(int64_t)m * nD + tid * P + u] = f;
        }
        FB_SUB(3)
        // lif.c phases 3-4 for the driven neurons
#pragma unroll
        for (int u = 0; u < DPT; u++) {
            Cell &c = cells[u];
            if (u < P && c.i >= 0) {
                membrane(d.k, c.st, c.bias);
                if (crosses(d.k, c.st, c.thr)) {
                    spike_effects(d.k, c.st, d.k.tRef);
                    c.count++;
                    atomicOr((unsigned long long *)&a.spk[(int64_t)m * a.nwords + (c.i >> 6)],
                             1ull << (c.i & 63));
                }
            }
        }
        FB_SUB(4)
        __syncthreads();   // s.draws is rewritten next step
        FB_SUB(5)
    }
#undef FB_SUB
}

// ------------------------------------------------------------------ E3: spike lists in lif.c order

__device__ void e3_list(const PersistArgs &a, Shared &s, int32_t done, int32_t m) {
    const DevBrain &d = a.d;
    const int tid = threadIdx.x;
    const int32_t nD = d.nDriven;
    const int32_t slot = (ring_slot(a, done, m) + d.nslots - 1) % d.nslots;
    int32_t *out = d.ring + (int64_t)slot * d.ringStride;
    // forced spikes in list order
    const int32_t P = (nD + PT - 1) / PT;
    const int32_t dlo = min(nD, tid * P), dhi = min(nD, dlo + P);
    const uint8_t *fl = a.dflag + (int64_t)m * nD;
    int32_t nf = 0;
    for (int32_t q = dlo; q < dhi; q++) nf += fl[q];
    int32_t ndTot;
    int32_t pos = wg_exclusive_scan<PT>(nf, s.scan, &ndTot);
    for (int32_t q = dlo; q < dhi; q++)
        if (fl[q]) out[pos++] = d.driven[q];
    // threshold spikes in ascending index
    uint64_t *bits = a.spk + (int64_t)m * a.nwords;
    const int32_t W = (a.nwords + PT - 1) / PT;
    const int32_t wlo = min(a.nwords, tid * W), whi = min(a.nwords, wlo + W);
    int32_t cnt = 0;
    for (int32_t w = wlo; w < whi; w++) cnt += __popcll(bits[w]);
    int32_t thTot;
    pos = ndTot + wg_exclusive_scan<PT>(cnt, s.scan, &thTot);
    for (int32_t w = wlo; w < whi; w++) {
        uint64_t b = bits[w];
        if (!b) continue;
        bits[w] = 0ull;
        while (b) {
            const int bit = __ffsll((unsigned long long)b) - 1;
            out[pos++] = (w << 6) + bit;
            b &= b - 1;
        }
    }
    if (tid == 0) {
        d.ringCount[slot] = ndTot + thTot;
        atomicAdd(&d.sc->spikes, (unsigned long long)(ndTot + thTot));
    }
}

// ------------------------------------------------------------------ the kernel

__global__ void __launch_bounds__(PT) k_persist(PersistArgs a) {
    __shared__ Shared s;
    cg::grid_group grid = cg::this_grid();
    const DevBrain &d = a.d;
    const int tid = threadIdx.x;
    const bool drivenBlock = blockIdx.x == 0;
    const int32_t lo = drivenBlock ? 0 : ((int32_t)blockIdx.x - 1) * a.nb;
    const int32_t hi = drivenBlock ? 0 : min(d.N, lo + a.nb);

    Cell cells[NPT];
    float pFire[DPT];
    uint32_t rng = 0;
    if (drivenBlock) {
        const int32_t P = (d.nDriven + PT - 1) / PT;
#pragma unroll
        for (int u = 0; u < DPT; u++) {
            const int32_t q = tid * P + u;
            const bool has = u < P && q < d.nDriven;
            load_cell(d, cells[u], has ? d.driven[q] : -1);
            pFire[u] = has ? d.drive[cells[u].i] * d.k.dtS : 0.f;
        }
        leap_load<PT>(s.leap, d.leap);
        rng = d.sc->rng;
    } else {
#pragma unroll
        for (int u = 0; u < NPT; u++) {
            const int32_t i = lo + tid + u * PT;
            load_cell(d, cells[u], (i < hi && a.drvPos[i] < 0) ? i : -1);
        }
    }

    // per-phase clock ticks, kept by thread 0 of blocks 0 and 1
    const bool timed = tid == 0 && blockIdx.x < 2;
    if (timed) {
        for (int k = 0; k < 6; k++) { s.ph[k] = 0; s.sub[k] = 0; }
        s.ph[6] = wall_clock64();
    }
#define FB_PHASE(k)                                    \
    if (timed) {                                       \
        const uint64_t now = wall_clock64();           \
        s.ph[k] += now - s.ph[6];                      \
        s.ph[6] = now;                                 \
    }
    int32_t done = 0;
    while (done < a.nsteps) {
        const int32_t L = min(a.L, a.nsteps - done);
        unsigned long long ev = 0, syn = 0;
        if (drivenBlock) e1_driven(a, s, done, L, ev, syn);
        else e1_regular(a, s, done, L, lo, hi);
        FB_PHASE(0)
        grid.sync();
        FB_PHASE(1)
        // every block reads the same flags after the barrier
        if (a.ctl[1] || a.ctl[2] > EV_CAP) break;
        if (ev) atomicAdd(&d.sc->edgeEvents, ev);
        if (syn) atomicAdd(&d.sc->synEvents, syn);
        if (drivenBlock) {
            e2_driven(a, s, cells, pFire, L, rng);
#pragma unroll
            for (int u = 0; u < DPT; u++) store_cell(d, cells[u]);
            if (tid == 0) d.sc->rng = rng;
        } else {
            e2_regular(a, s, cells, done, L, lo);
#pragma unroll
            for (int u = 0; u < NPT; u++) store_cell(d, cells[u]);
        }
        FB_PHASE(2)
        grid.sync();
        FB_PHASE(3)
        if ((int32_t)blockIdx.x < L) e3_list(a, s, done, (int32_t)blockIdx.x);
        if (blockIdx.x == gridDim.x - 1 && tid == 0) {
            // the slot delivered at the last step of the epoch is empty now
            d.ringCount[ring_slot(a, done, L - 1)] = 0;
            a.ctl[2] = 0;
        }
        FB_PHASE(4)
        grid.sync();
        FB_PHASE(5)
        done += L;
    }
#undef FB_PHASE
    if (timed) {
        for (int k = 0; k < 6; k++) atomicAdd((unsigned long long *)&a.prof[blockIdx.x * 6 + k], (unsigned long long)s.ph[k]);
        if (blockIdx.x == 0)
            for (int k = 0; k < 6; k++) atomicAdd((unsigned long long *)&a.prof[12 + k], (unsigned long long)s.sub[k]);
    }
    if (blockIdx.x == 0 && tid == 0) a.ctl[0] = done;
}

// ------------------------------------------------------------------ host side

struct PersistState {
    int device = -1;
    int maxBlocks = 0;
    double tickHz = 100e6;
};
static PersistState g_ps;

static fb_status persist_init() {
    int dev = 0;
    FB_HIP(hipGetDevice(&dev));
    if (g_ps.device == dev) return FB_OK;
    int perCU = 0;
    FB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k_persist, PT, 0));
    hipDeviceProp_t prop;
    FB_HIP(hipGetDeviceProperties(&prop, dev));
    if (!prop.cooperativeLaunch) {
        set_error("device does not support cooperative launch");
        return FB_ERR_UNSUPPORTED;
    }
    g_ps.maxBlocks = perCU * prop.multiProcessorCount;
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && khz > 0)
        g_ps.tickHz = (double)khz * 1000.0;
    g_ps.device = dev;
    return FB_OK;
}

// Smallest number of steps after a spike before the neuron can spike again.
static int32_t refractory_steps(const Consts &k) {
    float r = std::min(k.tRef, (k.tRef + k.dt) - k.dt);
    int32_t n = 0;
    while (r > 0 && n < 1000) { r = r - k.dt; n++; }
    return n;
}

fb_status persistent_supported(const DevBrain &d, bool serialDeliver, std::string *why) {
    if (serialDeliver) { *why = "presynaptic inhibition or tRef <= 0 needs the serial deliver path"; return FB_ERR_UNSUPPORTED; }
    if (d.k.bgEvents != 0.f) { *why = "background events draw from the shared stream every step"; return FB_ERR_UNSUPPORTED; }
    if (d.nDriven > DRV_CAP) { *why = "more than 2048 driven neurons"; return FB_ERR_UNSUPPORTED; }
    const int32_t D = d.nslots - 1;
    if (D < 1) { *why = "no delay"; return FB_ERR_UNSUPPORTED; }
    if (d.k.depU != 0.f && refractory_steps(d.k) < 1) {
        *why = "synaptic depression with a refractory period shorter than a step";
        return FB_ERR_UNSUPPORTED;
    }
    if (2LL * d.N >= (1LL << 20)) { *why = "more than 524,287 neurons (arrival index field)"; return FB_ERR_UNSUPPORTED; }
    if (d.g.maxRow >= (1LL << 26)) { *why = "a row longer than 2^26 edges (row position field)"; return FB_ERR_UNSUPPORTED; }
    fb_status st = persist_init();
    if (st) { *why = "cooperative launch unavailable"; return st; }
    const int32_t regular = std::max(1, g_ps.maxBlocks - 1);
    const int64_t nb = ((int64_t)(d.N + regular - 1) / regular + 63) / 64 * 64;
    if (nb > NB_MAX) { *why = "more than 2048 neurons per workgroup at this grid size"; return FB_ERR_UNSUPPORTED; }
    return FB_OK;
}

struct PersistBuffers {
    uint64_t *spk = nullptr;
    uint8_t *dflag = nullptr;
    int32_t *drvPos = nullptr;
    uint64_t *fwdKey = nullptr;
    float *fwdAdd = nullptr;
    int32_t *ctl = nullptr;
    uint64_t *prof = nullptr;
    int32_t N = 0, nDrivenCap = 0, fwdCap = 0;
    int32_t drivenVersion = -1;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

static std::map<const int32_t *, PersistBuffers> g_pbuf;   // keyed by the brain's driven buffer

fb_status launch_step(const DevBrain &d, int32_t head, int32_t slot, bool serialDeliver, bool rowsUnique,
                      hipStream_t st);

void persistent_release(const int32_t *key) {
    auto it = g_pbuf.find(key);
    if (it == g_pbuf.end()) return;
    PersistBuffers &pb = it->second;
    if (pb.spk) { hipFree(pb.spk); hipFree(pb.dflag); hipFree(pb.drvPos); hipFree(pb.fwdKey); hipFree(pb.fwdAdd); hipFree(pb.ctl); hipFree(pb.prof); }
    if (pb.ev0) { hipEventDestroy(pb.ev0); hipEventDestroy(pb.ev1); }
    g_pbuf.erase(it);
}

__global__ void k_drvpos(int32_t *drvPos, const int32_t *driven, int32_t N, int32_t nD) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) drvPos[i] = -1;
}
__global__ void k_drvpos_set(int32_t *drvPos, const int32_t *driven, int32_t nD) {
    const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nD) drvPos[driven[q]] = q;
}

fb_status launch_persistent(const DevBrain &d, int32_t head, int32_t nsteps, hipStream_t st, int64_t *fallbackSteps,
                            int64_t *launches, double *kernelMs, double *phaseUs) {
    std::string why;
    fb_status rc = persistent_supported(d, false, &why);
    if (rc) { set_error("persistent engine: %s", why.c_str()); return rc; }
    PersistBuffers &pb = g_pbuf[d.driven];
    const int32_t nwords = (d.N + 63) / 64;
    if (pb.N != d.N) {
        if (pb.spk) { hipFree(pb.spk); hipFree(pb.dflag); hipFree(pb.drvPos); hipFree(pb.fwdKey); hipFree(pb.fwdAdd); hipFree(pb.ctl); hipFree(pb.prof); }
        pb.N = d.N;
        pb.fwdCap = 1 << 16;
        FB_HIP(hipMalloc(&pb.spk, sizeof(uint64_t) * (size_t)D_MAX * nwords));
        FB_HIP(hipMemsetAsync(pb.spk, 0, sizeof(uint64_t) * (size_t)D_MAX * nwords, st));
        FB_HIP(hipMalloc(&pb.dflag, (size_t)D_MAX * DRV_CAP));
        FB_HIP(hipMalloc(&pb.drvPos, sizeof(int32_t) * (size_t)d.N));
        FB_HIP(hipMalloc(&pb.fwdKey, sizeof(uint64_t) * (size_t)pb.fwdCap));
        FB_HIP(hipMalloc(&pb.fwdAdd, sizeof(float) * (size_t)pb.fwdCap));
        FB_HIP(hipMalloc(&pb.ctl, sizeof(int32_t) * 4));
        FB_HIP(hipMalloc(&pb.prof, sizeof(uint64_t) * 18));
        if (!pb.ev0) {
            FB_HIP(hipEventCreate(&pb.ev0));
            FB_HIP(hipEventCreate(&pb.ev1));
        }
    }
    // driven positions (the list may have changed since the last launch)
    hipLaunchKernelGGL(k_drvpos, dim3((d.N + 255) / 256), dim3(256), 0, st, pb.drvPos, d.driven, d.N, d.nDriven);
    if (d.nDriven)
        hipLaunchKernelGGL(k_drvpos_set, dim3((d.nDriven + 255) / 256), dim3(256), 0, st, pb.drvPos, d.driven,
                           d.nDriven);
    FB_HIP(hipGetLastError());

    const int32_t D = d.nslots - 1;
    // With depression live, an epoch may not be longer than the refractory
    // period, so a neuron arrives at most once per epoch (E1 then derives its
    // res at arrival from the epoch-start value by recoveries alone).
    const int32_t L0 = std::min({D, d.k.depU != 0.f ? refractory_steps(d.k) : D, D_MAX});
    const int32_t regular = std::max(1, g_ps.maxBlocks - 1);
    const int32_t nb = (int32_t)(((int64_t)(d.N + regular - 1) / regular + 63) / 64 * 64);
    const int32_t need = (d.N + nb - 1) / nb;
    const int32_t grid = std::max(need + 1, std::min(g_ps.maxBlocks, std::max(L0, 2)));
    const int32_t L = std::min(L0, grid);   // block m materialises step m of an epoch

    PersistArgs a{};
    a.d = d;
    a.L = L;
    a.nb = nb;
    a.nwords = nwords;
    a.spk = pb.spk;
    a.dflag = pb.dflag;
    a.drvPos = pb.drvPos;
    a.fwdKey = pb.fwdKey;
    a.fwdAdd = pb.fwdAdd;
    a.fwdCap = pb.fwdCap;
    a.ctl = pb.ctl;
    a.prof = pb.prof;

    int32_t done = 0, cur = head;
    while (done < nsteps) {
        a.head0 = cur;
        a.nsteps = nsteps - done;
        FB_HIP(hipMemsetAsync(pb.ctl, 0, sizeof(int32_t) * 4, st));
        FB_HIP(hipMemsetAsync(pb.prof, 0, sizeof(uint64_t) * 18, st));
        void *args[] = {&a};
        FB_HIP(hipEventRecord(pb.ev0, st));
        FB_HIP(hipLaunchCooperativeKernel((const void *)k_persist, dim3(grid), dim3(PT), args, 0, st));
        FB_HIP(hipEventRecord(pb.ev1, st));
        if (launches) (*launches)++;
        int32_t ctl[4];
        FB_HIP(hipMemcpyAsync(ctl, pb.ctl, sizeof ctl, hipMemcpyDeviceToHost, st));
        FB_HIP(hipStreamSynchronize(st));
        float ms = 0.f;
        FB_HIP(hipEventElapsedTime(&ms, pb.ev0, pb.ev1));
        if (kernelMs) *kernelMs += ms;
        if (phaseUs) {
            uint64_t prof[18];
            FB_HIP(hipMemcpy(prof, pb.prof, sizeof prof, hipMemcpyDeviceToHost));
            for (int k = 0; k < 18; k++) phaseUs[k] += (double)prof[k] / g_ps.tickHz * 1e6;
        }
        done += ctl[0];
        cur = (cur + ctl[0]) % d.nslots;
        if (done < nsteps) {
            // an epoch overflowed a bounded buffer: run it with the per-step engine
            const int32_t m = std::min(L, nsteps - done);
            for (int32_t s2 = 0; s2 < m; s2++) {
                const int32_t slot = (cur + d.nslots - 1) % d.nslots;
                fb_status r2 = launch_step(d, cur, slot, false, true, st);
                if (r2) return r2;
                cur = (cur + 1) % d.nslots;
            }
            done += m;
            if (fallbackSteps) *fallbackSteps += m;
        }
    }
    return FB_OK;
}

}  // namespace fb
