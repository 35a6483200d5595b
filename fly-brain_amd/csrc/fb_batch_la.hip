// fb_batch_la.hip -- the ensemble lookahead engine: one cooperative launch
// advances all B members of an ensemble through many steps.
//
// As in the single-brain lookahead engine, a spike emitted at step t is
// delivered at t + D (D = nslots - 1), so while the grid simulates an epoch of
// D steps every arrival is already known, and one grid barrier per epoch is
// enough.  Lanes are members throughout.
//
//   regular blocks   block r owns neurons [80 r, 80 r + 80): wave w keeps 10
//                    neurons x 64 members in registers (lane = member).
//                    Driven neurons inside the range are not simulated here.
//   driven blocks    block g (g < ceil(B/4)) owns every driven neuron for the
//                    4 members 4g .. 4g+3.  Lane = (member m = lane & 3,
//                    slice j = lane >> 2); the driven list is cut into 128
//                    slices of at most 3 neurons per member.  Per step each
//                    lane counts its non-refractory neurons, a scan over the
//                    128 slices of the member (shuffles, then the 8 waves'
//                    totals in LDS) gives its offset in the member's serial
//                    xorshift32 stream, and a leapfrog product M^off x jumps
//                    there (M = one xorshift32 step over GF(2)).
//
// Per epoch, every block:
//   gather    loads the events addressed to it (one bucket per destination,
//             filled during the previous epoch), counting-sorts them into
//             cells (step, sign class, target) and sorts each cell by
//             (arrival rank, position in row) -- lif.c's addition order per
//             target -- then expands each event into one f32 addend per
//             member in its mask, W = f32(count * gain[type(pre)][member]) *
//             sign, so the step loop reads LDS only;
//   simulate  D steps from registers (regular blocks: no barrier between
//             steps; driven blocks: two per step for the stream scan);
//   push      walks the rows of its epoch's spikes twice: counts per
//             destination, one atomic per destination reserves a range of
//             that destination's bucket, then writes the events.  An event
//             is (cell, rank, position in row), the member mask (restricted
//             to the destination's members for a driven destination) and
//             (pre, synapse count).
// The arrival rank of a spike is its position class in lif.c's list: a forced
// spike's driven-list position, N + index for a threshold spike.
//
// The launch starts by pushing the spikes already in the delay line (from the
// per-step engine or the previous launch) and ends by writing its last
// epoch's spikes back into the delay line, so the per-step engine can continue
// at any point.  Preconditions (else the per-step engine runs): adaptation,
// depression and background off, D <= 18, at most 384 driven neurons, the
// grid fits the device.  A bucket overflow stops the launch at an epoch
// boundary; the host then runs that epoch on the per-step engine.

#include <algorithm>
#include <map>
#include <vector>

#include "fb_batch.h"


namespace fb {

constexpr int BT = 512;                    // threads per block
constexpr int BWV = BT / 64;               // waves per block
// A regular block's shape -- NPW neurons per wave, RNB = 8 NPW neurons per
// block -- is a template parameter: the autotuner picks it per ensemble
// (LA_SHAPES below).  Every shape gives the same bits: a block's shape decides
// which block simulates a neuron, never the order of any addition.
constexpr int MPG = 4;                     // members per driven block
constexpr int SPW = 64 / MPG;              // slices per wave
constexpr int NSL = SPW * BWV;             // slices of the driven list per member
constexpr int DPL = 3;                     // driven neurons per slice
constexpr int BDRV = NSL * DPL;            // driven neurons
constexpr int BDM = 18;                    // epoch length bound
constexpr int CELLS_D = BDM * 2 * BDRV;    // cells of a driven destination
constexpr int LEV = 2560;                  // events per destination per epoch
constexpr int LADD = 3072;                 // member addends per destination per epoch
constexpr int LSPK = 1024;                 // spike entries one block pushes per epoch
constexpr int LS_MAX = 512;                // destinations
constexpr int LMATS = 48;                  // M^1..M^32 and M^(32h), h < 16: offsets below 512
constexpr int LA_NPH = 5;                  // clocked phases: prologue, gather, simulate, push, barrier
constexpr uint32_t OWN_DRV = 0x80000000u;
constexpr int KEY_CELL_SHIFT = 47;

struct LaArgs {
    BatchDev d;
    int32_t t0, epochs, D, R, G, S, head0, nDrv, per;
    uint64_t *evKey, *evMask, *evAux;      // [2][S][LEV]
    int32_t *dstCnt;                       // [2][S] events reserved per destination
    int32_t *addCnt;                       // [2][S] member addends reserved per destination
    const uint32_t *eown;                  // [E] owner of each edge's target
    // [0] epochs done; [1] grid barrier timed out; [2 + p] overflow flags of
    // the pushes into parity p (written only while pushing, read only right
    // after a grid barrier)
    int32_t *ctl;
    int32_t *bar;                          // grid barrier counters (GBAR_WORDS, zero at launch)
    uint64_t *recM, *recF;                 // optional dense record [epochs*D][N]
    uint64_t *prof;                        // [LA_PROF] phase clocks (see k_batch_la), added per launch
    uint32_t *dspk;                        // [G][BDM][BDRV/4] each driven group's spike bytes of the epoch
    int32_t *dflag;                        // [G] 1 + the epoch whose bytes dspk holds (zero at launch)
};

template <int RNB>
struct LaShared {
    uint32_t cell[CELLS_D / 2];            // two u16 counters per word: cell ends after the gather
    uint64_t mask[LEV];
    uint16_t aoff[LEV];
    float add[LADD];
    union {
        struct {
            uint64_t key[LEV];
            uint64_t aux[LEV];
        } g;
        struct {
            int32_t pre[LSPK];
            uint32_t info[LSPK];           // step | forced << 8 | inhibitory << 9
            uint32_t rank[LSPK];           // arrival rank
            uint64_t m[LSPK];
            int32_t rowOff[LSPK + 1];
            uint32_t rowA[LSPK];
        } p;
    } u;
    union {
        uint64_t r[BDM][RNB];              // regular: threshold masks per (step, neuron)
        uint32_t d[BDM][BDRV / 4];         // driven: a byte per (step, neuron): forced | threshold << 4
    } spk;
    uint32_t leapS[LMATS * LEAP_STRIDE];
    uint64_t ph[LA_NPH + 1];               // phase clocks of thread 0, then the last stamp
    int32_t dcnt[LS_MAX];
    int32_t dbase[LS_MAX];
    int32_t dadd[LS_MAX];
    int32_t scan[64 + BWV];
    int32_t wtot[BWV][MPG];
    uint32_t x0[MPG];
    int32_t steps[64];
    uint64_t onm[BDM];                     // la_push: members running at each arrival step
    int32_t nEv, nSpk, nAdd;
};

template <int RNB>
__device__ __forceinline__ uint32_t cell_end(const LaShared<RNB> &s, int32_t c) {
    return (s.cell[c >> 1] >> ((c & 1) * 16)) & 0xFFFFu;
}

__device__ __forceinline__ uint32_t jump_to(const uint32_t *leapS, uint32_t x, int32_t off) {
    if (off <= 0) return x;
    const uint32_t y = matvec(leapS + (32 + (off >> 5)) * LEAP_STRIDE, x);
    return (off & 31) ? matvec(leapS + ((off & 31) - 1) * LEAP_STRIDE, y) : y;
}

// members running at absolute step t
template <int RNB>
__device__ __forceinline__ uint64_t on_mask(const LaShared<RNB> &s, int32_t B, int32_t t) {
    const int lane = threadIdx.x & 63;
    return __ballot(lane < B && t < s.steps[lane]);
}

// ------------------------------------------------------------------ gather

template <int RNB>
__device__ void la_gather(const LaArgs &a, LaShared<RNB> &s, int32_t par, int32_t me, bool drv) {
    const int tid = threadIdx.x;
    const int32_t CW = drv ? BDRV : RNB;
    const int32_t ncells = a.D * 2 * CW;
    const int32_t nw = (ncells + 1) >> 1;
    for (int32_t q = tid; q < nw; q += BT) s.cell[q] = 0;
    // more than LEV events (or LADD addends) were flagged when they were
    // pushed, and the launch stopped before this epoch
    const int32_t n = min(a.dstCnt[par * a.S + me], LEV);
    __syncthreads();
    const int64_t base = ((int64_t)par * a.S + me) * LEV;
    for (int32_t e = tid; e < n; e += BT) {
        const int32_t c = (int32_t)(a.evKey[base + e] >> KEY_CELL_SHIFT);
        atomicAdd(&s.cell[c >> 1], 1u << ((c & 1) * 16));
    }
    __syncthreads();
    // exclusive scan of the cell counts: each thread owns a run of whole
    // words (an even number of cells), so it rewrites them without atomics
    const int32_t per = (((ncells + BT - 1) / BT) + 1) & ~1;
    const int32_t c0 = min(2 * nw, tid * per), c1 = min(2 * nw, c0 + per);
    int32_t mine = 0;
    for (int32_t c = c0; c < c1; c++) mine += (int32_t)cell_end(s, c);
    int32_t total;
    int32_t run = wg_exclusive_scan<BT>(mine, s.scan, &total);
    for (int32_t c = c0; c < c1; c += 2) {
        const uint32_t lo = cell_end(s, c), hi = cell_end(s, c + 1);
        s.cell[c >> 1] = (uint32_t)run | ((uint32_t)(run + (int32_t)lo) << 16);
        run += (int32_t)(lo + hi);
    }
    __syncthreads();
    for (int32_t e = tid; e < n; e += BT) {
        const uint64_t k = a.evKey[base + e];
        const int32_t c = (int32_t)(k >> KEY_CELL_SHIFT);
        const uint32_t sh = (c & 1) * 16;
        const uint32_t pos = (atomicAdd(&s.cell[c >> 1], 1u << sh) >> sh) & 0xFFFFu;
        s.u.g.key[pos] = k & ((1ull << KEY_CELL_SHIFT) - 1);
        s.mask[pos] = a.evMask[base + e];
        s.u.g.aux[pos] = a.evAux[base + e];
    }
    __syncthreads();
    // each cell in (rank, position in row) order
    for (int32_t c = tid; c < ncells; c += BT) {
        const int32_t st = c ? (int32_t)cell_end(s, c - 1) : 0, en = (int32_t)cell_end(s, c);
        for (int32_t x = st + 1; x < en; x++) {
            const uint64_t kk = s.u.g.key[x], mm = s.mask[x], aa = s.u.g.aux[x];
            int32_t y = x;
            while (y > st && s.u.g.key[y - 1] > kk) {
                s.u.g.key[y] = s.u.g.key[y - 1];
                s.mask[y] = s.mask[y - 1];
                s.u.g.aux[y] = s.u.g.aux[y - 1];
                y--;
            }
            s.u.g.key[y] = kk;
            s.mask[y] = mm;
            s.u.g.aux[y] = aa;
        }
    }
    __syncthreads();
    // addend offsets: popcount prefix over the sorted events
    const int32_t pe = (n + BT - 1) / BT;
    const int32_t e0 = min(n, tid * pe), e1 = min(n, e0 + pe);
    int32_t cnt = 0;
    for (int32_t e = e0; e < e1; e++) cnt += __popcll(s.mask[e]);
    int32_t nadd;
    int32_t off = wg_exclusive_scan<BT>(cnt, s.scan, &nadd);
    for (int32_t e = e0; e < e1; e++) {
        s.aoff[e] = (uint16_t)min(off, LADD);
        off += __popcll(s.mask[e]);
    }
    (void)nadd;
    __syncthreads();
    // one addend per member in the mask: W = f32(count * gain) (writeGraph's
    // outScale rule, src/lifwasm.js:23,31) times the sign (res is 1)
    const BatchDev &d = a.d;
    for (int32_t e = tid; e < n; e += BT) {
        uint64_t m = s.mask[e];
        const int32_t pre = (int32_t)(s.u.g.aux[e] >> 32);
        const float cntf = (float)(uint32_t)(s.u.g.aux[e] & 0xFFFFFFFFu);
        const float sg = d.g.sign[pre];
        const float *gt = d.gainT + (int64_t)d.typeOf[pre] * d.B;
        int32_t o = s.aoff[e];
        while (m) {
            const int b = __ffsll((unsigned long long)m) - 1;
            m &= m - 1;
            if (o < LADD) s.add[o] = (cntf * gt[b]) * sg;
            o++;
        }
    }
    if (tid == 0) {
        a.dstCnt[par * a.S + me] = 0;   // this parity is pushed again only after the next barrier
        a.addCnt[par * a.S + me] = 0;
    }
    __syncthreads();
}

// the arrivals of one cell, in order, for this lane's member
template <int RNB>
__device__ __forceinline__ float apply_cell(const LaShared<RNB> &s, int32_t c, float g, int lane) {
    const int32_t st = c ? (int32_t)cell_end(s, c - 1) : 0, en = (int32_t)cell_end(s, c);
    const uint64_t below = (1ull << lane) - 1ull;
    for (int32_t k = st; k < en; k++) {
        const uint64_t m = s.mask[k];
        if ((m >> lane) & 1ull) {
            const int32_t o = s.aoff[k] + __popcll(m & below);
            if (o < LADD) g = g + s.add[o];
        }
    }
    return g;
}

// ------------------------------------------------------------------ push

// Appends a spike entry (neuron, step, forced?, mask) to the block's list.
// rank: the spike's position class in lif.c's list (driven position when it
// was emitted for a forced spike, N + index otherwise)
template <int RNB>
__device__ __forceinline__ void push_entry(const LaArgs &a, LaShared<RNB> &s, int32_t pre, int32_t step, bool forced,
                                           uint32_t rank, uint64_t m) {
    const float sg = a.d.g.sign[pre];
    if (sg == 0.f) return;                         // delivers nothing (lif.c:37)
    const int32_t k = atomicAdd(&s.nSpk, 1);
    if (k < LSPK) {
        s.u.p.pre[k] = pre;
        s.u.p.info[k] = (uint32_t)step | (forced ? 256u : 0u) | (sg > 0.f ? 0u : 512u);
        s.u.p.rank[k] = rank;
        s.u.p.m[k] = m;
    }
}

// destination of one edge event, or -1 when the mask has no bit left for it
__device__ __forceinline__ int32_t next_dst(const LaArgs &a, uint32_t own, uint64_t &mleft, uint64_t &mdst) {
    if (own & OWN_DRV) {
        if (!mleft) return -1;
        const int32_t g = (__ffsll((unsigned long long)mleft) - 1) / MPG;
        const uint64_t gm = 0xFull << (MPG * g);
        mdst = mleft & gm;
        mleft &= ~gm;
        return a.R + g;
    }
    if (!mleft) return -1;
    mdst = mleft;
    mleft = 0;
    return (int32_t)(own >> 8);
}

// Pushes the block's spike list into the buckets of parity `par`; the events
// arrive at absolute steps tArr0 + step.
template <int RNB>
__device__ void la_push(const LaArgs &a, LaShared<RNB> &s, int32_t par, int32_t tArr0) {
    const int tid = threadIdx.x;
    const BatchDev &d = a.d;
    __syncthreads();
    const int32_t ns = min(s.nSpk, LSPK);
    if (tid == 0 && s.nSpk > LSPK) atomicOr(&a.ctl[2 + par], 4);
    for (int32_t q = tid; q < a.S; q += BT) {
        s.dcnt[q] = 0;
        s.dadd[q] = 0;
    }
    // members running at each arrival step (one ballot per step)
    if (tid < 64)
        for (int32_t st = 0; st < a.D; st++) {
            const uint64_t on = on_mask(s, d.B, tArr0 + st);
            if (tid == 0) s.onm[st] = on;
        }
    __syncthreads();
    // Every loop below issues a batch of loads before it uses any: a loop
    // that used each global load before issuing the next paid a memory round
    // trip per iteration.
    constexpr int UB = 4;
    // row lengths (kept in rowOff until the scan) and the event counts
    const int32_t pe = (ns + BT - 1) / BT;
    const int32_t k0 = min(ns, tid * pe), k1 = min(ns, k0 + pe);
    int32_t len = 0;
    unsigned long long ev = 0, syn = 0;
    for (int32_t kb = k0; kb < k1; kb += UB) {
        uint32_t ra[UB], re[UB], sy[UB];
#pragma unroll
        for (int v = 0; v < UB; v++) {
            const int32_t pre = s.u.p.pre[min(kb + v, k1 - 1)];
            ra[v] = d.g.indptr[pre];
            re[v] = d.g.indptr[pre + 1];
            sy[v] = d.g.rowSyn[pre];
        }
#pragma unroll
        for (int v = 0; v < UB; v++) {
            const int32_t k = kb + v;
            if (k >= k1) break;
            s.u.p.rowA[k] = ra[v];
            s.u.p.rowOff[k] = (int32_t)(re[v] - ra[v]);
            len += (int32_t)(re[v] - ra[v]);
            const int32_t c = __popcll(s.u.p.m[k] & s.onm[s.u.p.info[k] & 255u]);
            ev += (unsigned long long)(re[v] - ra[v]) * (unsigned long long)c;
            syn += (unsigned long long)sy[v] * (unsigned long long)c;
        }
    }
    int32_t totalEdges;
    int32_t run = wg_exclusive_scan<BT>(len, s.scan, &totalEdges);
    for (int32_t k = k0; k < k1; k++) {
        const int32_t l = s.u.p.rowOff[k];
        s.u.p.rowOff[k] = run;
        run += l;
    }
    if (tid == 0) s.u.p.rowOff[ns] = totalEdges;
    if (ev) atomicAdd(&d.counters[0], ev);
    if (syn) atomicAdd(&d.counters[1], syn);
    __syncthreads();
    for (int pass = 0; pass < 2; pass++) {
        // thread tid takes edges tid, tid + BT, ... as before, UB at a time
        for (int32_t eb = tid; eb < totalEdges; eb += BT * UB) {
            int32_t kk[UB];
            uint32_t own[UB], cn[UB];
#pragma unroll
            for (int v = 0; v < UB; v++) {
                const int32_t e = min(eb + v * BT, totalEdges - 1);
                int32_t lo = 0, hi = ns - 1;           // last k with rowOff[k] <= e
                while (lo < hi) {
                    const int32_t mid = (lo + hi + 1) >> 1;
                    if (s.u.p.rowOff[mid] <= e) lo = mid; else hi = mid - 1;
                }
                kk[v] = lo;
                const uint32_t j = s.u.p.rowA[lo] + (uint32_t)(e - s.u.p.rowOff[lo]);
                own[v] = a.eown[j];
                cn[v] = pass ? (uint32_t)d.g.counts[j] : 0u;
            }
#pragma unroll
            for (int v = 0; v < UB; v++) {
                const int32_t e = eb + v * BT;
                if (e >= totalEdges) break;
                const int32_t k = kk[v];
                const uint32_t jr = (uint32_t)(e - s.u.p.rowOff[k]);
                uint64_t mleft = s.u.p.m[k], mdst = 0;
                int32_t dst;
                while ((dst = next_dst(a, own[v], mleft, mdst)) >= 0) {
                    if (pass == 0) {
                        atomicAdd(&s.dcnt[dst], 1);
                        atomicAdd(&s.dadd[dst], __popcll(mdst));
                        continue;
                    }
                    const int32_t pos = s.dbase[dst] + atomicAdd(&s.dcnt[dst], 1);
                    if (pos >= LEV) continue;              // flagged at the reservation
                    const int32_t pre = s.u.p.pre[k];
                    const uint32_t info = s.u.p.info[k];
                    const int32_t step = (int32_t)(info & 255u);
                    const uint32_t cls = (info >> 9) & 1u;
                    const bool drvDst = (own[v] & OWN_DRV) != 0;
                    const uint32_t tloc = drvDst ? (own[v] & 0xFFFFu) : (own[v] & 0xFFu);
                    const uint32_t CW = drvDst ? (uint32_t)BDRV : (uint32_t)RNB;
                    const uint64_t cell = ((uint64_t)step * 2u + cls) * CW + tloc;
                    const uint64_t rank = s.u.p.rank[k];
                    const int64_t at = ((int64_t)par * a.S + dst) * LEV + pos;
                    a.evKey[at] = (cell << KEY_CELL_SHIFT) | (rank << 23) | (uint64_t)jr;
                    a.evMask[at] = mdst;
                    a.evAux[at] = ((uint64_t)(uint32_t)pre << 32) | (uint64_t)cn[v];
                }
            }
        }
        __syncthreads();
        if (pass == 0) {
            for (int32_t q = tid; q < a.S; q += BT) {
                const int32_t c = s.dcnt[q];
                if (c) {
                    const int32_t b0 = atomicAdd(&a.dstCnt[par * a.S + q], c);
                    s.dbase[q] = b0;
                    if (b0 + c > LEV) atomicOr(&a.ctl[2 + par], 2);
                    if (atomicAdd(&a.addCnt[par * a.S + q], s.dadd[q]) + s.dadd[q] > LADD)
                        atomicOr(&a.ctl[2 + par], 1);
                }
                s.dcnt[q] = 0;
            }
            __syncthreads();
        }
    }
    if (tid == 0) s.nSpk = 0;
    __syncthreads();
}

// The delay line's pending spikes (arrivals at t0 .. t0 + D - 1): block `me`
// takes every nblk-th entry of the D slot lists.
template <int RNB>
__device__ void la_prologue(const LaArgs &a, LaShared<RNB> &s, int32_t me, int32_t nblk) {
    const BatchDev &d = a.d;
    if (threadIdx.x == 0) s.nSpk = 0;
    __syncthreads();
    int32_t off[BDM + 1];
    off[0] = 0;
    for (int32_t st = 0; st < a.D; st++) off[st + 1] = off[st] + d.nActive[(a.head0 + st) % d.nslots];
    const int32_t tot = off[a.D];
    // entries g = me, me + nblk, me + 2 nblk, ... of the D lists back to back
    for (int32_t k = threadIdx.x;; k += BT) {
        const int64_t g = me + (int64_t)nblk * k;
        if (g >= tot) break;
        int32_t st = 0;
        while (off[st + 1] <= g) st++;
        const int32_t slot = (a.head0 + st) % d.nslots;
        const int32_t i = d.active[(int64_t)slot * d.N + (g - off[st])];
        const uint64_t m = d.mask[(int64_t)slot * d.N + i], f = d.fmask[(int64_t)slot * d.N + i];
        if (f) push_entry(a, s, i, st, true, (uint32_t)d.frank[(int64_t)slot * d.N + i], f);
        if (m & ~f) push_entry(a, s, i, st, false, (uint32_t)d.N + (uint32_t)i, m & ~f);
    }
    la_push(a, s, 0, a.t0);
}

// ------------------------------------------------------------------ the kernel

template <int NPW>
__global__ void __launch_bounds__(BT) k_batch_la(LaArgs a) {
    constexpr int RNB = BWV * NPW;
    static_assert(sizeof(LaShared<RNB>) <= 160 * 1024, "LDS of one workgroup");
    __shared__ LaShared<RNB> s;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int32_t me = blockIdx.x;
    const BatchDev &d = a.d;
    const int32_t N = d.N, B = d.B, D = a.D;
    const bool drv = me < a.G;
    const int32_t dstId = drv ? a.R + me : me - a.G;
    for (int32_t q = tid; q < LMATS * LEAP_WORDS; q += BT) s.leapS[(q >> 7) * LEAP_STRIDE + (q & 127)] = d.leap[q];
    if (tid < 64) s.steps[tid] = tid < B ? d.steps[tid] : 0;
    if (tid == 0) { s.nSpk = 0; s.nEv = 0; s.nAdd = 0; }
    __syncthreads();

    // ---- registers: regular neurons (lane = member) or this lane's driven slice
    float v[NPW], gE[NPW], gI[NPW], rf[NPW], tr[NPW], th[NPW], bi[NPW];
    uint32_t sc[NPW];
    bool valid[NPW];
    const int32_t rbase = drv ? 0 : (me - a.G) * RNB;
    // driven role: member m of group g, slice jj
    const int m4 = lane & (MPG - 1);
    const int32_t bm = drv ? me * MPG + m4 : lane;
    const int32_t slice = wave * SPW + (lane >> 2);
    float pf[DPL];
#pragma unroll
    for (int n = 0; n < NPW; n++) {
        int32_t i = -1;
        if (!drv) {
            const int32_t q = rbase + wave * NPW + n;
            i = (q < N && d.drvPos[q] < 0) ? q : -1;
        } else if (n < DPL) {
            const int32_t p = slice * a.per + n;
            i = (n < a.per && p < a.nDrv) ? d.driven[p] : -1;
        }
        valid[n] = i >= 0;
        const bool ld = valid[n] && bm < B;
        const int64_t ib = (int64_t)(i < 0 ? 0 : i) * B + (bm < B ? bm : 0);
        v[n] = ld ? d.v[ib] : 0.f;
        gE[n] = ld ? d.gE[ib] : 0.f;
        gI[n] = ld ? d.gI[ib] : 0.f;
        rf[n] = ld ? d.refr[ib] : 0.f;
        tr[n] = ld ? d.trace[ib] : 0.f;
        th[n] = ld ? d.thr[ib] : 0.f;
        bi[n] = ld ? d.bias[ib] : 0.f;
        sc[n] = ld ? d.spikeCount[ib] : 0u;
        if (drv && n < DPL) pf[n] = valid[n] ? d.drive[i] * d.k.dtS : 0.f;
    }
    if (drv && tid < MPG) s.x0[tid] = me * MPG + tid < B ? d.rng[me * MPG + tid] : 0u;

    // clocks of the first driven and the first regular workgroup, per phase:
    // prologue, gather, simulate, push, grid barrier
    const int timedSlot = me == 0 ? 0 : me == a.G ? 1 : -1;
    const bool timed = tid == 0 && timedSlot >= 0;
    if (timed) {
        for (int k = 0; k < LA_NPH; k++) s.ph[k] = 0;
        s.ph[LA_NPH] = wall_clock64();
    }
#define LA_PHASE(k)                                    \
    if (timed) {                                       \
        const uint64_t now = wall_clock64();           \
        s.ph[k] += now - s.ph[LA_NPH];                 \
        s.ph[LA_NPH] = now;                            \
    }
    la_prologue(a, s, me, gridDim.x);
    LA_PHASE(0)
    if (!grid_barrier(a.bar, &a.ctl[1])) return;
    LA_PHASE(4)

    unsigned long long spikes = 0;
    int32_t last = -1;
    // an overflow in the prologue (pushed into parity 0) leaves the first
    // epoch's events incomplete: nothing is simulated and the delay line stays
    // as it was
    for (int32_t e = 0; e < a.epochs && !(e == 0 && a.ctl[2]); e++) {
        const int32_t par = e & 1;
        la_gather(a, s, par, dstId, drv);
        LA_PHASE(1)
        const int32_t tE = a.t0 + e * D;
        if (!drv) {
            for (int32_t st = 0; st < D; st++) {
                const bool on = lane < B && tE + st < s.steps[lane];
#pragma unroll
                for (int n = 0; n < NPW; n++) {
                    if (!valid[n]) {
                        if (lane == 0) s.spk.r[st][wave * NPW + n] = 0ull;
                        continue;
                    }
                    const int32_t tl = wave * NPW + n;
                    bool spk = false;
                    if (on) {
                        NeuronState ns;
                        ns.gE = apply_cell(s, (st * 2 + 0) * RNB + tl, gE[n], lane);
                        ns.gI = apply_cell(s, (st * 2 + 1) * RNB + tl, gI[n], lane);
                        ns.v = v[n]; ns.refr = rf[n]; ns.trace = tr[n]; ns.adapt = 0.f; ns.res = 1.f;
                        membrane(d.k, ns, bi[n]);
                        if (crosses(d.k, ns, th[n])) {
                            spike_effects(d.k, ns, d.k.tRef);
                            sc[n] += 1u;
                            spk = true;
                        }
                        v[n] = ns.v; gE[n] = ns.gE; gI[n] = ns.gI; rf[n] = ns.refr; tr[n] = ns.trace;
                    }
                    const uint64_t mk = __ballot(spk);
                    if (lane == 0) {
                        s.spk.r[st][tl] = mk;
                        spikes += (unsigned long long)__popcll(mk);
                        if (mk && a.recM) a.recM[(int64_t)(e * D + st) * N + rbase + tl] = mk;
                    }
                }
            }
        } else {
            for (int32_t q = tid; q < D * (BDRV / 4); q += BT) s.spk.d[q / (BDRV / 4)][q % (BDRV / 4)] = 0u;
            __syncthreads();
            for (int32_t st = 0; st < D; st++) {
                const bool on = bm < B && tE + st < s.steps[bm];
                const uint32_t xs0 = s.x0[m4];
                int32_t cnt = 0;
#pragma unroll
                for (int n = 0; n < DPL; n++) {
                    if (valid[n] && on) {
                        const int32_t p = slice * a.per + n;
                        gE[n] = apply_cell(s, (st * 2 + 0) * BDRV + p, gE[n], bm);
                        gI[n] = apply_cell(s, (st * 2 + 1) * BDRV + p, gI[n], bm);
                        if (!(rf[n] > 0)) cnt++;
                    }
                }
                // position in the member's stream: scan over its 128 slices
                int32_t incl = cnt;
#pragma unroll
                for (int o = MPG; o < 64; o <<= 1) {
                    const int32_t y = __shfl_up(incl, o, 64);
                    if (lane >= o) incl += y;
                }
                if ((lane >> 2) == SPW - 1) s.wtot[wave][m4] = incl;
                __syncthreads();
                int32_t off = incl - cnt, total = 0;
                for (int w = 0; w < BWV; w++) {
                    const int32_t t_ = s.wtot[w][m4];
                    off += w < wave ? t_ : 0;
                    total += t_;
                }
                uint32_t x = cnt > 0 ? jump_to(s.leapS, xs0, off) : xs0;
                bool forced[DPL];
#pragma unroll
                for (int n = 0; n < DPL; n++) {
                    forced[n] = false;
                    if (valid[n] && on && !(rf[n] > 0)) {
                        x = xs32(x);
                        const float u = (float)(x >> 8) * (1.0f / 16777216.0f);
                        forced[n] = u < pf[n];
                    }
                }
#pragma unroll
                for (int n = 0; n < DPL; n++) {
                    if (!(valid[n] && on)) continue;
                    NeuronState ns;
                    ns.v = v[n]; ns.gE = gE[n]; ns.gI = gI[n]; ns.refr = rf[n]; ns.trace = tr[n];
                    ns.adapt = 0.f; ns.res = 1.f;
                    if (forced[n]) {
                        // lif.c:51-52, then the membrane pass sees the neuron refractory
                        ns.v = d.k.vReset;
                        ns.refr = d.k.tRef + d.k.dt;
                        ns.trace = 1.f;
                        sc[n] += 1u;
                    }
                    membrane(d.k, ns, bi[n]);
                    bool thr = false;
                    if (crosses(d.k, ns, th[n])) {
                        spike_effects(d.k, ns, d.k.tRef);
                        sc[n] += 1u;
                        thr = true;
                    }
                    v[n] = ns.v; gE[n] = ns.gE; gI[n] = ns.gI; rf[n] = ns.refr; tr[n] = ns.trace;
                    if (forced[n] || thr) {
                        const int32_t p = slice * a.per + n;
                        const uint32_t byte = (forced[n] ? 1u << m4 : 0u) | (thr ? 16u << m4 : 0u);
                        atomicOr(&s.spk.d[st][p >> 2], byte << (8 * (p & 3)));
                        spikes++;
                        if (a.recM) {
                            const int64_t at = (int64_t)(e * D + st) * N + d.driven[p];
                            atomicOr((unsigned long long *)&a.recM[at], 1ull << bm);
                            if (forced[n]) atomicOr((unsigned long long *)&a.recF[at], 1ull << bm);
                        }
                    }
                }
                // the stream after every draw of this step: the last slice's lane
                if (slice == NSL - 1) s.x0[m4] = cnt > 0 ? x : jump_to(s.leapS, xs0, total);
                __syncthreads();
            }
        }
        // push this epoch's spikes for the next one.  A driven group's rows
        // are most of an epoch's events and its workgroup finishes its steps
        // first, so it publishes its spike bytes and every workgroup pushes a
        // share of one group's rows beside its own
        if (e + 1 < a.epochs && drv) {
            __syncthreads();
            uint32_t *out = a.dspk + (int64_t)me * BDM * (BDRV / 4);
            for (int32_t q = tid; q < D * (BDRV / 4); q += BT) out[q] = s.spk.d[q / (BDRV / 4)][q % (BDRV / 4)];
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            if (tid == 0) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                __hip_atomic_store(&a.dflag[me], e + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        LA_PHASE(2)
        if (e + 1 < a.epochs) {
            __syncthreads();
            if (!drv) {
                for (int32_t q = tid; q < D * RNB; q += BT) {
                    const int32_t st = q / RNB, tl = q % RNB;
                    const uint64_t mk = s.spk.r[st][tl];
                    if (mk) push_entry(a, s, rbase + tl, st, false, (uint32_t)N + (uint32_t)(rbase + tl), mk);
                }
            }
            // workgroup k takes driven group g = k mod G and, of its (step,
            // position) pairs q, those with q = k div G (mod H), H the
            // workgroups of that group
            const int32_t g = me % a.G, jg = me / a.G, H = (a.S - g + a.G - 1) / a.G;
            if (tid == 0) {
                const uint64_t t0 = wall_clock64();
                while (__hip_atomic_load(&a.dflag[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < e + 1) {
                    __builtin_amdgcn_s_sleep(1);
                    if (wall_clock64() - t0 > GBAR_TIMEOUT) {
                        atomicOr(&a.ctl[1], GBAR_TIMED_OUT);
                        break;
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
            __syncthreads();
            const uint32_t *src = a.dspk + (int64_t)g * BDM * (BDRV / 4);
            const uint64_t sh = (uint64_t)MPG * g;
            for (int32_t q = jg + tid * H; q < D * a.nDrv; q += BT * H) {
                const int32_t st = q / a.nDrv, p = q % a.nDrv;
                const uint32_t byte = (src[st * (BDRV / 4) + (p >> 2)] >> (8 * (p & 3))) & 255u;
                if (!byte) continue;
                const int32_t i = d.driven[p];
                if (byte & 15u) push_entry(a, s, i, st, true, (uint32_t)p, (uint64_t)(byte & 15u) << sh);
                if (byte >> 4) push_entry(a, s, i, st, false, (uint32_t)N + (uint32_t)i, (uint64_t)(byte >> 4) << sh);
            }
            la_push(a, s, (e + 1) & 1, tE + D);
        }
        LA_PHASE(3)
        last = e;
        if (!grid_barrier(a.bar, &a.ctl[1])) return;
        LA_PHASE(4)
        // the pushes of this epoch went to parity (e + 1) & 1; the next writes
        // to that flag come two epochs later, after every block has read it
        if (e + 1 < a.epochs && a.ctl[2 + ((e + 1) & 1)]) break;
    }
    if (spikes) atomicAdd(&d.counters[2], spikes);
    if (timed && a.prof)
        for (int k = 0; k < LA_NPH; k++) atomicAdd((unsigned long long *)&a.prof[timedSlot * LA_NPH + k], (unsigned long long)s.ph[k]);
#undef LA_PHASE

    // ---- write back: state, streams, and the last epoch's spikes into the delay line
#pragma unroll
    for (int n = 0; n < NPW; n++) {
        if (!valid[n] || bm >= B) continue;
        int32_t i;
        if (!drv) i = rbase + wave * NPW + n;
        else i = d.driven[slice * a.per + n];
        const int64_t ib = (int64_t)i * B + bm;
        d.v[ib] = v[n]; d.gE[ib] = gE[n]; d.gI[ib] = gI[n]; d.refr[ib] = rf[n]; d.trace[ib] = tr[n];
        d.spikeCount[ib] = sc[n];
    }
    if (drv && tid < MPG && me * MPG + tid < B) d.rng[me * MPG + tid] = s.x0[tid];
    if (me == 0 && tid == 0) a.ctl[0] = last + 1;
    if (last < 0) return;                              // uniform over the grid
    const int32_t tL = last * D;                       // first step of the last epoch, launch-relative
    // clear the slots the last epoch fills (all blocks), then fill them
    for (int32_t st = 0; st < D; st++) {
        const int32_t slot = (a.head0 + tL + st + d.nslots - 1) % d.nslots;
        if (!drv) {
            for (int32_t q = tid; q < RNB; q += BT) {
                const int32_t i = rbase + q;
                if (i < N) { d.mask[(int64_t)slot * N + i] = 0ull; d.fmask[(int64_t)slot * N + i] = 0ull; }
            }
        } else if (me == 0) {
            for (int32_t p = tid; p < a.nDrv; p += BT) {
                const int32_t i = d.driven[p];
                d.mask[(int64_t)slot * N + i] = 0ull;
                d.fmask[(int64_t)slot * N + i] = 0ull;
            }
        }
    }
    if (me == 0)
        for (int32_t q = tid; q < d.nslots; q += BT) d.nActive[q] = 0;
    if (!grid_barrier(a.bar, &a.ctl[1])) return;
    for (int32_t st = 0; st < D; st++) {
        const int32_t slot = (a.head0 + tL + st + d.nslots - 1) % d.nslots;
        if (!drv) {
            for (int32_t q = tid; q < RNB; q += BT) {
                const uint64_t mk = s.spk.r[st][q];
                if (!mk) continue;
                const int32_t i = rbase + q;
                d.mask[(int64_t)slot * N + i] = mk;
                const int32_t pos = atomicAdd(&d.nActive[slot], 1);
                d.active[(int64_t)slot * N + pos] = i;
            }
        } else {
            for (int32_t p = tid; p < a.nDrv; p += BT) {
                const uint32_t byte = (s.spk.d[st][p >> 2] >> (8 * (p & 3))) & 255u;
                if (!byte) continue;
                const int32_t i = d.driven[p];
                const uint64_t sh = (uint64_t)MPG * me;
                const uint64_t fb = (uint64_t)(byte & 15u) << sh, tb = (uint64_t)(byte >> 4) << sh;
                const unsigned long long old = atomicOr((unsigned long long *)&d.mask[(int64_t)slot * N + i],
                                                        (unsigned long long)(fb | tb));
                if (fb) {
                    atomicOr((unsigned long long *)&d.fmask[(int64_t)slot * N + i], (unsigned long long)fb);
                    d.frank[(int64_t)slot * N + i] = p;
                }
                if (!old) {
                    const int32_t pos = atomicAdd(&d.nActive[slot], 1);
                    d.active[(int64_t)slot * N + pos] = i;
                }
            }
        }
    }
}

// owner of each edge's target: a driven neuron's list position (with the
// driven bit), else regular block << 8 | position in the block
__global__ void k_batch_eown(const uint2 *edges, int64_t E, const int32_t *drvPos, uint32_t *eown, int32_t rnb) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < E; j += (int64_t)gridDim.x * blockDim.x) {
        const int32_t q = (int32_t)edges[j].x;
        const int32_t dp = drvPos[q];
        eown[j] = dp >= 0 ? (OWN_DRV | (uint32_t)dp) : ((uint32_t)(q / rnb) << 8) | (uint32_t)(q % rnb);
    }
}

// ------------------------------------------------------------------ host side

// The shapes the autotuner chooses among: neurons per wave of a regular block
// (72, 80, 96 or 128 neurons per block).  Fewer, larger regular blocks mean fewer
// destinations to push to and gather from per epoch, and a longer serial walk
// per wave per step; which wins depends on the graph, the activity and the
// device, so the engine measures it.
constexpr int LA_SHAPES[LA_NSHAPES] = {9, 10, 12, 16};

static const void *la_kernel(int k) {
    switch (k) {
        case 0: return (const void *)k_batch_la<9>;
        case 1: return (const void *)k_batch_la<10>;
        case 2: return (const void *)k_batch_la<12>;
        default: return (const void *)k_batch_la<16>;
    }
}

int32_t batch_la_shape_neurons(int k) { return BWV * LA_SHAPES[k]; }

struct LaBuffers {
    uint64_t *prof = nullptr;                    // [2 * LA_NPH] phase clocks of the last launch
    uint32_t *dspk = nullptr;                    // [LS_MAX / MPG][BDM][BDRV / 4]
    int32_t *dflag = nullptr;                    // [LS_MAX / MPG]
    uint64_t *evKey = nullptr, *evMask = nullptr, *evAux = nullptr;
    int32_t *dstCnt = nullptr, *ctl = nullptr;   // dstCnt holds [2][S] events, then [2][S] addends
    int32_t *bar = nullptr;
    uint32_t *eown = nullptr;
    int32_t S = 0;                               // destinations the buffers hold (grown, never shrunk)
    int64_t E = 0;
    int32_t eownRnb = 0;
    uint64_t drivenVersion = ~0ull;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

static std::map<const void *, LaBuffers> g_labuf;    // keyed by the ensemble's device state

static int32_t la_max_blocks(int k) {
    static int32_t maxB[LA_NSHAPES] = {-1, -1, -1, -1};
    if (maxB[k] < 0) {
        int dev = 0, ncu = 0, per = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, la_kernel(k), BT, 0) != hipSuccess) return 0;
        maxB[k] = ncu * per;
    }
    return maxB[k];
}

static bool la_common(const BatchDev &d, std::string *why) {
    const int32_t D = d.nslots - 1;
    auto no = [&](const char *w) { if (why) *why = w; return false; };
    if (d.liveAdapt || d.liveRes || d.liveBg) return no("adaptation, depression or background is on");
    if (d.signGain || d.laneDrive || d.poolOf) return no("lane mode or the pool loop is on");
    if (D < 1 || D > BDM) return no("delay outside 1..18 steps");
    if (d.nDriven > BDRV) return no("more than 384 driven neurons");
    if (d.g.maxRow >= (1 << 23)) return no("a row longer than 2^23");
    if ((int64_t)2 * d.N >= (1 << 24)) return no("N too large for the arrival rank");
    return true;
}

bool batch_la_shape_fits(const BatchDev &d, int k) {
    const int32_t rnb = batch_la_shape_neurons(k);
    const int32_t R = (d.N + rnb - 1) / rnb, G = (d.B + MPG - 1) / MPG;
    if (R + G > LS_MAX) return false;
    const int32_t maxB = la_max_blocks(k);
    return R + G <= maxB && R + G <= GBAR_MAX_BLOCKS;
}

bool batch_la_supported(const BatchDev &d, std::string *why) {
    if (!la_common(d, why)) return false;
    for (int k = 0; k < LA_NSHAPES; k++)
        if (batch_la_shape_fits(d, k)) return true;
    if (why) *why = "the grid does not fit the device";
    return false;
}

void batch_la_release(const void *key) {
    std::lock_guard<std::mutex> lock(grid_launch_mutex());
    auto it = g_labuf.find(key);
    if (it == g_labuf.end()) return;
    LaBuffers &b = it->second;
    for (void *p : {(void *)b.evKey, (void *)b.evMask, (void *)b.evAux, (void *)b.dstCnt, (void *)b.ctl,
                    (void *)b.eown, (void *)b.bar, (void *)b.prof, (void *)b.dspk, (void *)b.dflag})
        if (p) (void)hipFree(p);
    if (b.ev0) (void)hipEventDestroy(b.ev0);
    if (b.ev1) (void)hipEventDestroy(b.ev1);
    g_labuf.erase(it);
}

// Runs `epochs` whole epochs from absolute step t0 with ring head head0, with
// regular blocks of shape k.  Returns the epochs actually done in *done (fewer
// after a bucket overflow) and adds the kernel's HIP-event time to *kernelMs.
fb_status batch_la_launch(const BatchDev &d, int k, int32_t t0, int32_t head0, int32_t epochs,
                          uint64_t drivenVersion, uint64_t *recM, uint64_t *recF, hipStream_t st, int32_t *done,
                          double *kernelMs, double *phaseUs) {
    std::lock_guard<std::mutex> lock(grid_launch_mutex());
    const int32_t D = d.nslots - 1;
    const int32_t rnb = batch_la_shape_neurons(k);
    const int32_t R = (d.N + rnb - 1) / rnb, G = (d.B + MPG - 1) / MPG, S = R + G;
    LaBuffers &b = g_labuf[(const void *)d.v];
    if (b.S < S) {
        for (void *p : {(void *)b.evKey, (void *)b.evMask, (void *)b.evAux, (void *)b.dstCnt})
            if (p) FB_HIP(hipFree(p));
        const size_t n = 2 * (size_t)S * LEV;
        FB_HIP(hipMalloc(&b.evKey, n * 8));
        FB_HIP(hipMalloc(&b.evMask, n * 8));
        FB_HIP(hipMalloc(&b.evAux, n * 8));
        FB_HIP(hipMalloc(&b.dstCnt, 4 * sizeof(int32_t) * S));
        b.S = S;
    }
    if (!b.ctl) {
        FB_HIP(hipMalloc(&b.ctl, 4 * sizeof(int32_t)));
        FB_HIP(hipMalloc(&b.bar, sizeof(int32_t) * GBAR_WORDS));
        FB_HIP(hipMalloc(&b.prof, sizeof(uint64_t) * 2 * LA_NPH));
        FB_HIP(hipMalloc(&b.dspk, sizeof(uint32_t) * (LS_MAX / MPG) * BDM * (BDRV / 4)));
        FB_HIP(hipMalloc(&b.dflag, sizeof(int32_t) * (LS_MAX / MPG)));
        FB_HIP(hipEventCreate(&b.ev0));
        FB_HIP(hipEventCreate(&b.ev1));
    }
    if (b.E != d.g.E || b.drivenVersion != drivenVersion || b.eownRnb != rnb) {
        if (b.E != d.g.E || !b.eown) {
            if (b.eown) FB_HIP(hipFree(b.eown));
            FB_HIP(hipMalloc(&b.eown, sizeof(uint32_t) * std::max<int64_t>(1, d.g.E)));
        }
        const int grid = (int)std::min<int64_t>(4096, (d.g.E + 255) / 256 + 1);
        hipLaunchKernelGGL(k_batch_eown, dim3(grid), dim3(256), 0, st, d.g.edges, d.g.E, d.drvPos, b.eown, rnb);
        FB_HIP(hipGetLastError());
        b.E = d.g.E;
        b.drivenVersion = drivenVersion;
        b.eownRnb = rnb;
    }
    FB_HIP(hipMemsetAsync(b.dstCnt, 0, 4 * sizeof(int32_t) * S, st));
    FB_HIP(hipMemsetAsync(b.ctl, 0, 4 * sizeof(int32_t), st));
    FB_HIP(hipMemsetAsync(b.bar, 0, sizeof(int32_t) * GBAR_WORDS, st));
    FB_HIP(hipMemsetAsync(b.prof, 0, sizeof(uint64_t) * 2 * LA_NPH, st));
    FB_HIP(hipMemsetAsync(b.dflag, 0, sizeof(int32_t) * (LS_MAX / MPG), st));
    LaArgs a;
    a.d = d;
    a.t0 = t0; a.epochs = epochs; a.D = D; a.R = R; a.G = G; a.S = S; a.head0 = head0;
    a.nDrv = d.nDriven;
    a.per = std::max(1, (d.nDriven + NSL - 1) / NSL);
    a.evKey = b.evKey; a.evMask = b.evMask; a.evAux = b.evAux; a.dstCnt = b.dstCnt; a.addCnt = b.dstCnt + 2 * S;
    a.eown = b.eown; a.ctl = b.ctl; a.bar = b.bar; a.recM = recM; a.recF = recF; a.prof = b.prof;
    a.dspk = b.dspk; a.dflag = b.dflag;
    void *args[] = {&a};
    FB_HIP(hipEventRecord(b.ev0, st));
    const hipError_t le = hipLaunchCooperativeKernel(la_kernel(k), dim3(S), dim3(BT), args, 0, st);
    if (le != hipSuccess) {
        // refused (the resource behind grid-wide barriers is held elsewhere): nothing ran
        (void)hipGetLastError();
        set_error("ensemble lookahead: cooperative launch refused: %s", hipGetErrorString(le));
        return FB_ERR_UNSUPPORTED;
    }
    FB_HIP(hipEventRecord(b.ev1, st));
    int32_t ctl[4];
    uint64_t prof[2 * LA_NPH];
    FB_HIP(hipMemcpyAsync(ctl, b.ctl, sizeof ctl, hipMemcpyDeviceToHost, st));
    FB_HIP(hipMemcpyAsync(prof, b.prof, sizeof prof, hipMemcpyDeviceToHost, st));
    FB_HIP(hipStreamSynchronize(st));
    float ms = 0.f;
    FB_HIP(hipEventElapsedTime(&ms, b.ev0, b.ev1));
    if (kernelMs) *kernelMs += ms;
    if (phaseUs) {
        static double tickHz = 0.0;
        if (tickHz == 0.0) {
            int dev = 0, khz = 0;
            tickHz = hipGetDevice(&dev) == hipSuccess &&
                             hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && khz > 0
                         ? khz * 1000.0
                         : 1e8;
        }
        for (int k = 0; k < 2 * LA_NPH; k++) phaseUs[k] += (double)prof[k] / tickHz * 1e6;
    }
    if (ctl[1] & GBAR_TIMED_OUT) {
        set_error("ensemble lookahead: a grid barrier timed out; the ensemble state is undefined");
        return FB_ERR_DEVICE;
    }
    *done = ctl[0];
    return FB_OK;
}

}  // namespace fb
