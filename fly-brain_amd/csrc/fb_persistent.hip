// fb_persistent.hip -- the lookahead engine: one grid-barrier launch advances
// many LIF steps, with the delay line as a lookahead window.
//
// A spike emitted at step t is delivered at step t + D (D = nslots - 1, 18 at
// dt 0.1 ms).  The launch advances in epochs of D steps.  Every arrival of
// epoch e+1 is a spike of epoch e, so while the grid simulates epoch e no
// neuron needs anything from another workgroup -- except the forced spikes,
// which draw from one serial xorshift32 stream in driven-list order.  Roles:
//
//   blocks 0..G-1   the driven blocks: block k owns the driven-list positions
//                   [kQ, (k+1)Q) (Q <= 2048; state in registers, list order
//                   across threads).  Per step: a ballot scan gives each
//                   thread its position in the stream, and every thread jumps
//                   to its first draw with nibble-table products and takes
//                   the later ones by xorshift steps (M = one xorshift32 step
//                   as a matrix over GF(2)).  With G > 1 the blocks swap their
//                   non-refractory counts through tagged words in global
//                   memory (xc): block k needs the counts of blocks < k for
//                   its stream offset and every count of the previous step
//                   for the stream state.  A neuron's refractoriness at step
//                   t is fixed by the end of step t-1 (arrivals do not touch
//                   it), so each count is published one step ahead, and the
//                   swap stays off the serial chain.
//   blocks G..R     regular blocks: block b owns a contiguous, 64-aligned
//                   range of up to 1024 neurons (2 per thread, in registers).
//   blocks R+1..    spare blocks (when the grid has them): they turn the
//                   previous epoch's spikes into ring lists in lif.c order.
//
// Within epoch e every owning block:
//   gather  loads the edge events addressed to it (pushed during epoch e-1)
//           into LDS and sorts them by (step, sign class, target, arrival
//           key, position in row) -- lif.c's addition order per target;
//   simulate  D steps from registers: arrivals (one add, or the sorted run
//           for several), forced spikes, membrane, threshold.  Threshold
//           spikes go to a per-step bitmask; the driven block keeps its
//           neurons' spike steps as bit masks and writes them once per epoch
//           (a global store inside its step loop would make every workgroup
//           barrier wait for the store to complete);
//   push    walks the rows of its own spikes and writes their events into
//           the destinations' buckets for epoch e+1 (a slot is one atomic on
//           the destination's count).  Once every regular block has pushed
//           and the driven blocks have written their spike masks, the 64
//           blocks after the driven ones push the driven rows the same way,
//           and one grid barrier ends the epoch.
// The arrival key of a spike is its rank in lif.c's list order: a forced
// spike's driven-list position, or 8192 + index for a threshold spike.
//
// A launch starts with a prologue in which the regular blocks scatter the
// ring lists already in the delay line (from the per-step engines or the
// previous launch) into the buckets, keyed by list position, while the
// driven blocks load their cells from the stage (DSTAGE_F).  Preconditions (checked on the host;
// otherwise the per-step engines run): no background events, no presynaptic
// inhibition, D <= 32, at most 8192 driven neurons and 1024 neurons per
// block, and with synaptic depression live a refractory period of at least D
// steps (a neuron then arrives at most once per epoch).  A bucket or LDS
// overflow stops the launch at an epoch boundary; the host runs one epoch on
// the per-step engine and relaunches.

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "fb_internal.h"


namespace fb {

// 512 threads per block: the kernel may then use 256 VGPRs per lane, which
// keeps the driven block's per-step state out of scratch memory.
constexpr int PT = 512;             // threads per block
constexpr int NPT = 2;              // neurons per thread in a regular block
constexpr int NB_MAX = PT * NPT;    // neurons per regular block
constexpr int D_MAX = 32;           // epoch length bound
constexpr int EV_CAP = 6144;        // events per destination block per epoch (LDS)
constexpr int BK_STRIDE = 32;       // words between two destinations' bucket counts (their own 128-byte lines)
constexpr int SPK_CAP = 2048;       // spikes per block per epoch (LDS push list)
constexpr int RPT = SPK_CAP / PT;   // push rows per thread in the row scan
constexpr int DPT = 4;              // driven neurons per thread in a driven block
constexpr int DRV_BLK = PT * DPT;   // driven neurons per driven block
constexpr int DRV_BLOCKS = 4;       // driven blocks at most
constexpr int DRV_CAP = DRV_BLK * DRV_BLOCKS;
constexpr int XC_RING = 8;          // exchange words per driven block (a ring over steps)
constexpr int CELLS = DPT > NPT ? DPT : NPT;
constexpr int S_MAX = 512;          // owning blocks (driven + regular)
constexpr int WPS = (NB_MAX > DRV_BLK ? NB_MAX : DRV_BLK) / 32;   // words of one step in the hasEv bitmap
constexpr uint32_t KEY_THRESHOLD = DRV_CAP;   // arrival key offset of threshold spikes
constexpr int LEAP_PERSIST_MATS = LEAP_TABLE_MATS;   // M^1..M^32, M^(32h), and M^2048, all in LDS
#ifndef FB_SHARE_BLOCKS
#define FB_SHARE_BLOCKS 64
#endif
constexpr int SHARE_BLOCKS = FB_SHARE_BLOCKS;   // blocks that push the driven block's rows
constexpr int SHQ = 256;            // driven neurons per share block
constexpr int LONG_ROW = 1024;      // a regular block's rows longer than this are pushed by the whole grid
constexpr int LR_CAP = SPK_CAP;     // long rows per epoch in the shared list (more stay with their block)
constexpr int RUN_LIST = 1024;      // (step, target) runs a gather can list for ordering by a wave
constexpr int LR_WORDS = 64 + GBAR_WORDS + DRV_BLOCKS;   // lrCnt[2], pflag[2], then from word 64 epc
// PersistArgs::flags: the stage already holds the driven neurons' state (the
// last launch left it there) / their bias, thr, sign, drive and index
constexpr int32_t STAGE_STATE = 2, STAGE_PARAMS = 4;

// event key: step(5) | local target(12) | class(1) | arrival key(20) | position in row(26).
// Sorted, the events of one (step, target) are contiguous, excitatory first,
// each class in lif.c's addition order.
__device__ __forceinline__ uint64_t ev_key(uint32_t m, uint32_t c, uint32_t ql, uint32_t k, uint32_t jr) {
    return ((uint64_t)m << 59) | ((uint64_t)ql << 47) | ((uint64_t)c << 46) | ((uint64_t)k << 26) | (uint64_t)jr;
}
__device__ __forceinline__ uint32_t ev_slot(uint64_t key) { return (uint32_t)(key >> 47); }   // step << 12 | target
__device__ __forceinline__ uint32_t ev_class(uint64_t key) { return (uint32_t)(key >> 46) & 1u; }

struct PersistArgs {
    DevBrain d;
    int32_t head0;       // ring head at launch
    int32_t nsteps;
    int32_t D;           // epoch length (the delay in steps)
    int32_t nb;          // largest regular block, in neurons (a multiple of 64, at most NB_MAX)
    const int32_t *bstart;   // [S-G+1] regular block r owns neurons bstart[r] .. bstart[r+1]-1
    int32_t G;           // driven blocks (blocks 0 .. G-1)
    int32_t Q;           // driven-list positions per driven block
    int32_t S;           // owning blocks: the driven blocks + regular blocks
    int32_t nSpare;      // spare blocks after them
    int32_t nwords;      // 64-bit words of one step's spike bitmask
    uint64_t *spk;       // [2][D_MAX][nwords] threshold spikes of non-driven neurons
    uint32_t *dmask;     // [2][DRV_CAP] steps with a forced spike, by driven-list position
    uint32_t *tmask;     // [2][DRV_CAP] steps with a threshold spike of a driven neuron
    const int32_t *drvPos;   // [N] position in the driven list, -1 if not driven
    uint32_t *arrMask;   // [N] arrival steps of the first epoch (depression bookkeeping)
    uint64_t *bkKey;     // [2][S][EV_CAP] each destination's bucket of the next epoch's events
    float *bkAdd;
    int32_t *bkCnt;      // [2][S] slots taken, one per BK_STRIDE words (zero at launch; a destination clears its own after reading)
    int32_t *ctl;        // [0] steps done, [1] overflow flags, [2] blocks that have left (zero at launch)
    int32_t *hostCtl;    // pinned host memory: [0] steps done, [1] flags, [2] tag, from the last block out
    uint64_t *hostProf;  // pinned host memory: the phase clocks, from the last block out
    int32_t tag;         // this launch's number
    float *dstage;       // [DSTAGE_F][DRV_CAP] the driven neurons' state at launch, in driven-list order
    uint64_t *prof;      // phase clocks (see k_persist)
    const uint32_t *eown;    // [E] destination block << 12 | local index of each edge's target
    float *dres;         // [2][DRV_CAP] res of each driven neuron at the epoch's end
    int32_t *bar;        // grid barrier counters (GBAR_WORDS, zero at launch)
    uint64_t *xc;        // [DRV_BLOCKS][XC_RING] (step + 1) << 32 | non-refractory count, zero at launch
    int32_t *lrCnt;      // [2] long rows queued by the epochs of each parity (zero at launch)
    int32_t *pflag;      // [2] overflow flags of the pushes into the buckets of each parity (lrCnt + 2)
    int32_t *epc;        // regular workgroups whose pushes are done, summed over the launch's epochs and
                         // counted as grid_barrier counts arrivals: [64 (g + 1)] the regular workgroups of
                         // group g (GBAR_GROUP a group), [0] the groups all of whose regular workgroups
                         // have; [GBAR_WORDS + k] 1 + the last epoch whose masks driven workgroup k wrote
                         // (lrCnt + 64; zero at launch)
    int32_t *lrI;        // [2][LR_CAP] a queued long row: neuron, step << 24 | arrival key, sign x res at arrival
    uint32_t *lrK;
    float *lrR;
    int32_t flags;       // bit 0: time the driven block's step parts; STAGE_STATE, STAGE_PARAMS
};

struct Shared {
    // The epoch's events live from the gather to the end of the simulation,
    // the push list from the push to the next gather: they share the space.
    union {
        struct {
            uint64_t evKey[EV_CAP];
            float evAdd[EV_CAP];
            uint16_t evStart[EV_CAP];    // the r-th (step, target) with arrivals: counts, then starts, then ends of its run
        };
        struct {
            int32_t spkI[SPK_CAP];       // push list: neuron, step << 24 | arrival key, sign x res at arrival
            uint32_t spkK[SPK_CAP];
            float spkR[SPK_CAP];
            int32_t rowOff[SPK_CAP + 1]; // first edge of each pushed row in the block's flattened edge list
            uint32_t rowA[SPK_CAP];      // row start in the CSR
        };
    };
    uint32_t hasEv[D_MAX * WPS];     // bit (step, target): the target has arrivals at that step
    int32_t evRank[D_MAX * WPS];     // set bits of hasEv before each word
    uint32_t leap[LEAP_PERSIST_MATS * LEAP_STRIDE];
    int32_t pref[D_MAX + 1];     // prologue: first ring entry of each arrival step
    int32_t scan[64 + PT / 64];
    uint32_t rngNext[2];
    int32_t nev, nspk, last;
    int32_t nRunList;                // sort_runs: runs ordered by a wave
    uint16_t runList[RUN_LIST];
    unsigned long long ev, syn;
    uint64_t ph[NPH + 1];        // phase clocks of thread 0, then the last stamp
    uint64_t sub[7];             // driven-step clocks of thread 0
    uint64_t share[4];           // shared-push clocks of thread 0: load, edges, publish, last stamp
    // this block's share of the driven neurons (positions q0 .. q0+nq-1), fixed per launch
    int32_t shI[SHQ];            // neuron
    uint32_t shRA[SHQ];          // row start in the CSR
    int32_t shLen[SHQ];          // row length
    uint32_t shSyn[SHQ];         // synapses of the row
    float shSign[SHQ];
    int32_t shQ0, shNq;
};

// ------------------------------------------------------------------ helpers

// The run of a (step, target) pair: its rank among the pairs with events of
// the epoch, counted in hasEv bit order (step-major, then target), which is
// the order of the sorted keys.
__device__ __forceinline__ int32_t run_of(const Shared &s, uint32_t m, uint32_t ql) {
    const int32_t w = (int32_t)(m * WPS + (ql >> 5));
    return s.evRank[w] + __popc(s.hasEv[w] & ((1u << (ql & 31)) - 1u));
}
__device__ __forceinline__ int32_t run_of_key(const Shared &s, uint64_t key) {
    const uint32_t sl = ev_slot(key);
    return run_of(s, sl >> 12, sl & 4095u);
}

// evStart as 32-bit words for LDS atomics: run r is the 16-bit half r & 1 of
// word r >> 1 (every value stays below EV_CAP < 65,536, so no carry crosses
// halves).  Adds 1 to run r's half and returns its old value.
__device__ __forceinline__ uint32_t run_inc(Shared &s, int32_t r) {
    uint32_t *w = reinterpret_cast<uint32_t *>(s.evStart);
    const uint32_t sh = (uint32_t)(r & 1) * 16u;
    return (atomicAdd(&w[r >> 1], 1u << sh) >> sh) & 0xFFFFu;
}

// Runs of at most RUN_SMALL events are insertion-sorted by one thread; longer
// ones are listed (up to RUN_LIST of them) and ordered by a wave.
constexpr int RUN_SMALL = 4;
constexpr int GREG = 4;             // events per thread a gather ranks in registers

// After the counting sort every run r = [end(r-1), end(r)) (end = evStart)
// holds its events in arbitrary order: order each by key (class, arrival
// key, position in row), which is lif.c's addition order.  A short run is
// insertion-sorted in place by one thread.  A longer run goes to a list, and
// a wave ranks it: each lane takes one event (a run of up to 64) and
// counts the smaller keys, and the lanes write their events to their ranks
// once every lane has read the run -- or, for a run longer than 64, the
// lanes write to this block's bucket (free once read) and the run is copied
// back after a block barrier.  More than RUN_LIST long runs: the waves walk
// every run instead of the list.  Every thread of the block calls it.
__device__ __forceinline__ void sort_runs(const PersistArgs &a, Shared &s, int32_t nRuns, int64_t base) {
    const int tid = threadIdx.x;
    if (tid == 0) s.nRunList = 0;
    __syncthreads();
    int big = 0;
    for (int32_t r = tid; r < nRuns; r += PT) {
        const int32_t b0 = r ? s.evStart[r - 1] : 0, b1 = s.evStart[r];
        if (b1 - b0 < 2) continue;
        if (b1 - b0 > RUN_SMALL) {
            const int32_t at = atomicAdd(&s.nRunList, 1);
            if (at < RUN_LIST) s.runList[at] = (uint16_t)r;
            big |= b1 - b0 > 64;
            continue;
        }
        for (int32_t i = b0 + 1; i < b1; i++) {
            const uint64_t k = s.evKey[i];
            const float v = s.evAdd[i];
            int32_t j = i - 1;
            while (j >= b0 && s.evKey[j] > k) {
                s.evKey[j + 1] = s.evKey[j];
                s.evAdd[j + 1] = s.evAdd[j];
                j--;
            }
            s.evKey[j + 1] = k;
            s.evAdd[j + 1] = v;
        }
    }
    big = __syncthreads_or(big);
    const int32_t nList = s.nRunList;
    if (nList == 0) return;
    const bool listed = nList <= RUN_LIST;
    const int32_t count = listed ? nList : nRuns;
    const int lane = tid & 63, wave = tid >> 6;
    for (int32_t x = wave; x < count; x += PT / 64) {
        const int32_t r = listed ? (int32_t)s.runList[x] : x;
        const int32_t b0 = r ? s.evStart[r - 1] : 0, b1 = s.evStart[r];
        if (b1 - b0 <= RUN_SMALL) continue;                 // uniform in the wave
        if (b1 - b0 <= 64) {
            const int32_t q = b0 + lane;
            uint64_t k = 0;
            float v = 0.f;
            int32_t rank = 0;
            if (q < b1) { k = s.evKey[q]; v = s.evAdd[q]; }
            for (int32_t h = b0; h < b1; h++) rank += s.evKey[h] < k;   // keys are distinct
            __builtin_amdgcn_wave_barrier();
            if (q < b1) { s.evKey[b0 + rank] = k; s.evAdd[b0 + rank] = v; }
        } else {
            for (int32_t q = b0 + lane; q < b1; q += 64) {
                const uint64_t k = s.evKey[q];
                int32_t rank = 0;
                for (int32_t h = b0; h < b1; h++) rank += s.evKey[h] < k;
                a.bkKey[base + b0 + rank] = k;
                a.bkAdd[base + b0 + rank] = s.evAdd[q];
            }
        }
    }
    if (!big) { __syncthreads(); return; }
    __threadfence_block();
    __syncthreads();
    for (int32_t x = wave; x < count; x += PT / 64) {
        const int32_t r = listed ? (int32_t)s.runList[x] : x;
        const int32_t b0 = r ? s.evStart[r - 1] : 0, b1 = s.evStart[r];
        if (b1 - b0 <= 64) continue;
        for (int32_t q = b0 + lane; q < b1; q += 64) {
            s.evKey[q] = a.bkKey[base + q];
            s.evAdd[q] = a.bkAdd[base + q];
        }
    }
    __syncthreads();
}

// This step's arrivals on one neuron: its (step, target) run, excitatory
// events first, each class in lif.c's order
__device__ __forceinline__ void apply_run(const Shared &s, uint32_t m, uint32_t ql, float &gE, float &gI) {
    const int32_t r = run_of(s, m, ql);
    const int32_t b0 = r ? s.evStart[r - 1] : 0, b1 = s.evStart[r];
    for (int32_t q = b0; q < b1; q++) {
        if (ev_class(s.evKey[q])) gI = gI + s.evAdd[q]; else gE = gE + s.evAdd[q];
    }
}

// res at the start of epoch step m from its value at the epoch start: m recoveries
__device__ __forceinline__ float res_after(const Consts &k, float r, int32_t m) {
    for (int32_t j = 0; j < m; j++) r = r + (1.f - r) * k.kRec;
    return r;
}

// ring head at step `t` of the launch
__device__ __forceinline__ int32_t head_at(const PersistArgs &a, int32_t t) { return (a.head0 + t) % a.d.nslots; }

__device__ __forceinline__ int64_t bucket_at(const PersistArgs &a, int32_t par, int32_t dst) {
    return ((int64_t)par * a.S + dst) * EV_CAP;
}

// Row lengths of the `ns` rows of the LDS push list -> the flattened edge
// list: s.rowOff[r] (first edge of row r, s.rowOff[ns] = total) and s.rowA[r]
// (row start in the CSR).  Row r is (s.spkI[r], step and key in s.spkK[r],
// sign x res in s.spkR[r]); a row whose sign x res is 0 adds nothing.  With
// `count`, adds the rows' edge and synapse counts to s.ev / s.syn.  With
// lpar >= 0, a row longer than LONG_ROW goes to the shared list of parity
// lpar instead (while it has room) and adds nothing here: once every regular
// block has pushed, every block pushes a slice of those rows (push_long_share).
// Returns the total.
__device__ __forceinline__ int32_t row_scan(const PersistArgs &a, Shared &s, int32_t ns, bool count = true,
                                   int32_t lpar = -1) {
    const DevBrain &d = a.d;
    const int tid = threadIdx.x;
    // row lengths -> prefix (RPT consecutive rows per thread)
    int32_t len[RPT];
    uint32_t ra[RPT];
    int32_t mySum = 0, evs = 0;
    unsigned long long syn = 0;
#pragma unroll
    for (int v = 0; v < RPT; v++) {
        const int32_t r = RPT * tid + v;
        len[v] = 0;
        ra[v] = 0;
        if (r < ns && s.spkR[r] != 0.f) {
            const int32_t i = s.spkI[r];
            ra[v] = d.g.indptr[i];
            len[v] = (int32_t)(d.g.indptr[i + 1] - ra[v]);
            syn += d.g.rowSyn[i];   // 0 for an empty row
        }
        evs += len[v];
        if (lpar >= 0 && len[v] > LONG_ROW) {
            const int32_t q = atomicAdd(&a.lrCnt[lpar], 1);
            if (q < LR_CAP) {
                const int64_t at = (int64_t)lpar * LR_CAP + q;
                a.lrI[at] = s.spkI[r];
                a.lrK[at] = s.spkK[r];
                a.lrR[at] = s.spkR[r];
                len[v] = 0;
            }
        }
        mySum += len[v];
    }
    int32_t total;
    int32_t off = wg_exclusive_scan<PT>(mySum, s.scan, &total);
#pragma unroll
    for (int v = 0; v < RPT; v++) {
        const int32_t r = RPT * tid + v;
        if (r < ns) { s.rowOff[r] = off; s.rowA[r] = ra[v]; }
        off += len[v];
    }
    if (tid == 0) s.rowOff[ns] = total;
    if (count && evs) { atomicAdd(&s.ev, (unsigned long long)evs); atomicAdd(&s.syn, syn); }
    __syncthreads();
    return total;
}

// An overflow of a push into the buckets of parity par.  The host reads the
// flags in ctl[1]; the blocks read pflag[par] at the top of the epoch that
// gathers those buckets, a word no block writes during that epoch (its pushes
// go to the other parity), so every block takes the same exit.
__device__ __forceinline__ void flag_overflow(const PersistArgs &a, int32_t par, int32_t bit) {
    atomicOr(&a.ctl[1], bit);
    atomicOr(&a.pflag[par], bit);
}

// Writes the edge events g0 <= g < g1 of the flattened edge list (row_scan)
// into the destinations' buckets for parity `par`, four edges per thread in
// flight; a scan of the row starts maps an edge to its row.  A slot is one
// atomic on the destination's count; beyond EV_CAP the event is dropped and
// flag 2 stops the launch before the epoch that would read it.
__device__ __forceinline__ void push_edges(const PersistArgs &a, Shared &s, int32_t par, int32_t ns, int32_t g0, int32_t g1) {
    const DevBrain &d = a.d;
    const int tid = threadIdx.x;
    int32_t *cnt = a.bkCnt + (int64_t)par * a.S * BK_STRIDE;
    constexpr int U = 4;
    for (int32_t gb = g0 + tid; gb < g1; gb += PT * U) {
        uint2 ed[U];
        uint32_t own[U];
        int32_t row[U];
        uint32_t j[U];
#pragma unroll
        for (int v = 0; v < U; v++) {
            const int32_t g = gb + v * PT;
            row[v] = -1;
            if (g < g1) {
                int32_t lo = 0, hi = ns;   // last row with rowOff <= g
                while (hi - lo > 1) {
                    const int32_t mid = (lo + hi) >> 1;
                    if (s.rowOff[mid] <= g) lo = mid; else hi = mid;
                }
                row[v] = lo;
                j[v] = s.rowA[lo] + (uint32_t)(g - s.rowOff[lo]);
                ed[v] = d.g.edges[j[v]];
                own[v] = a.eown[j[v]];
            }
        }
#pragma unroll
        for (int v = 0; v < U; v++) {
            if (row[v] < 0) continue;
            const int32_t r = row[v];
            const uint32_t kk = s.spkK[r];
            const float sv = s.spkR[r];
            // a +0 weight adds -0 to gI, which changes nothing (x + -0 = x),
            // and +0 to gE, which changes nothing while no excitatory
            // conductance can be -0 (skipZeroW)
            if (ed[v].y == 0u && (sv < 0.f || d.skipZeroW)) continue;
            const int32_t dst = (int32_t)(own[v] >> 12);
            const uint32_t ql = own[v] & 4095u;
            const int32_t pos = atomicAdd(&cnt[(int64_t)dst * BK_STRIDE], 1);
            if (pos < EV_CAP) {
                const int64_t at = bucket_at(a, par, dst) + pos;
                a.bkKey[at] = ev_key(kk >> 24, sv > 0 ? 0u : 1u, ql, kk & 0xFFFFFFu, j[v] - s.rowA[r]);
                a.bkAdd[at] = __uint_as_float(ed[v].y) * sv;
            } else if (pos == EV_CAP) {
                flag_overflow(a, par, 2);
            }
        }
    }
}

// Writes the events of the `ns` rows of the LDS push list into the buckets
// for parity `par`.
__device__ __forceinline__ void push_rows(const PersistArgs &a, Shared &s, int32_t par, int32_t ns, int32_t lpar = -1) {
    const int32_t total = row_scan(a, s, ns, true, lpar);
    push_edges(a, s, par, ns, 0, total);
}

// The long rows the regular blocks queued this epoch (list parity lpar), as
// one flattened edge list split evenly over the grid: block b writes edges
// [b T / G, (b+1) T / G) into the buckets of parity par.  The owners counted
// their events.  Every thread of every block calls it.
__device__ __forceinline__ void push_long_share(const PersistArgs &a, Shared &s, int32_t par, int32_t lpar) {
    const int tid = threadIdx.x;
    const int32_t n = min(__hip_atomic_load(&a.lrCnt[lpar], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), LR_CAP);
    if (n == 0) return;   // the same value in every block once every regular block has pushed
    __syncthreads();      // the LDS push list may still be read by this block's driven share push
    for (int32_t r = tid; r < n; r += PT) {
        const int64_t at = (int64_t)lpar * LR_CAP + r;
        s.spkI[r] = a.lrI[at];
        s.spkK[r] = a.lrK[at];
        s.spkR[r] = a.lrR[at];
    }
    __syncthreads();
    const int32_t total = row_scan(a, s, n, false);
    const int64_t G = gridDim.x, b = blockIdx.x;
    push_edges(a, s, par, n, (int32_t)(total * b / G), (int32_t)(total * (b + 1) / G));
    __syncthreads();
}

// Folds this block's work counters (row_scan) into the brain's.
__device__ __forceinline__ void publish_stats(const PersistArgs &a, Shared &s) {
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s.ev) atomicAdd(&a.d.sc->edgeEvents, s.ev);
        if (s.syn) atomicAdd(&a.d.sc->synEvents, s.syn);
        s.ev = 0;
        s.syn = 0;
    }
}

// ------------------------------------------------------------------ prologue: ring lists -> buckets

__device__ __forceinline__ void prologue_scatter(const PersistArgs &a, Shared &s, int32_t L0) {
    const DevBrain &d = a.d;
    const int tid = threadIdx.x;
    const int32_t me = blockIdx.x;
    // the arrival steps' list lengths, loaded in parallel, then their prefix
    if (tid < L0) s.pref[tid + 1] = d.ringCount[head_at(a, tid)];
    __syncthreads();
    if (tid == 0) {
        int32_t acc = 0;
        for (int32_t m = 0; m < L0; m++) { const int32_t c = s.pref[m + 1]; s.pref[m] = acc; acc += c; }
        s.pref[L0] = acc;
    }
    __syncthreads();
    const int32_t Atot = s.pref[L0];
    // regular block r of R scatters arrivals g = r, r + R, r + 2R, ... in
    // chunks of SPK_CAP rows (not the driven blocks: their prologue is the
    // launch's critical path)
    const int32_t rb = me - a.G, R = a.S - a.G;
    const int32_t mine = Atot > rb ? (Atot - rb + R - 1) / R : 0;
    for (int32_t c0 = 0; c0 < mine; c0 += SPK_CAP) {
        const int32_t ns = min(SPK_CAP, mine - c0);
        for (int32_t r = tid; r < ns; r += PT) {
            const int32_t g = rb + (c0 + r) * R;
            int32_t m = 0;
            while (s.pref[m + 1] <= g) m++;
            const int32_t k = g - s.pref[m];
            const int32_t pre = d.ring[(int64_t)head_at(a, m) * d.ringStride + k];
            // a driven neuron's res is the stage's while the stage holds its state
            float res = d.res[pre];
            if (a.flags & STAGE_STATE) {
                const int32_t q = a.drvPos[pre];
                if (q >= 0) res = a.dstage[6 * DRV_CAP + q];
            }
            const float sv = d.g.sign[pre] * res_after(d.k, res, m);
            if (sv != 0.f && d.k.depU != 0.f) atomicOr(&a.arrMask[pre], 1u << m);
            s.spkI[r] = pre;
            s.spkK[r] = ((uint32_t)m << 24) | (uint32_t)k;
            s.spkR[r] = sv;
        }
        __syncthreads();
        push_rows(a, s, 0, ns);
        __syncthreads();
    }
    publish_stats(a, s);
}

// ------------------------------------------------------------------ gather: bucket -> sorted LDS events

// The epoch's events of this block, from its bucket into LDS, ordered by
// (step, target, class, arrival key, position in row): a counting sort over
// the (step, target) pairs with events -- which pairs, how many events each,
// where each goes -- then each pair's run in key order (ranks within the
// run, or sort_runs past GREG * PT events).  Afterwards evStart[r] is the
// end of run r.
__device__ __forceinline__ int32_t gather(const PersistArgs &a, Shared &s, int32_t par, int32_t me) {
    const int tid = threadIdx.x;
    // the count is read once, then cleared for the push two epochs on
    if (tid == 0) {
        int32_t *cnt = a.bkCnt + ((int64_t)par * a.S + me) * BK_STRIDE;
        s.nev = min(*cnt, EV_CAP);
        *cnt = 0;
    }
    for (int32_t w = tid; w < D_MAX * WPS; w += PT) s.hasEv[w] = 0u;
    uint32_t *rw = reinterpret_cast<uint32_t *>(s.evStart);
    for (int32_t w = tid; w < EV_CAP / 2; w += PT) rw[w] = 0u;
    __syncthreads();
    const int32_t n = s.nev;
    const int64_t base = bucket_at(a, par, me);
    // which (step, target) pairs have events (the keys copied to LDS for the count)
    for (int32_t g = tid; g < n; g += PT) {
        const uint64_t key = a.bkKey[base + g];
        s.evKey[g] = key;
        const uint32_t sl = ev_slot(key);
        atomicOr(&s.hasEv[(sl >> 12) * WPS + ((sl & 4095u) >> 5)], 1u << (sl & 31u));
    }
    __syncthreads();
    // each pair's run index: set bits of hasEv before it
    constexpr int W = D_MAX * WPS, PER = W / PT;
    int32_t c = 0;
#pragma unroll
    for (int v = 0; v < PER; v++) c += __popc(s.hasEv[tid * PER + v]);
    int32_t nRuns;
    int32_t off = wg_exclusive_scan<PT>(c, s.scan, &nRuns);
#pragma unroll
    for (int v = 0; v < PER; v++) { s.evRank[tid * PER + v] = off; off += __popc(s.hasEv[tid * PER + v]); }
    __syncthreads();
    // events per run
    for (int32_t g = tid; g < n; g += PT) run_inc(s, run_of_key(s, s.evKey[g]));
    __syncthreads();
    // counts -> first slot of each run (16-bit stores: two runs share a word)
    const int32_t per = (nRuns + PT - 1) / PT;
    const int32_t r0 = min(nRuns, tid * per), r1 = min(nRuns, r0 + per);
    int32_t sum = 0;
    for (int32_t r = r0; r < r1; r++) sum += s.evStart[r];
    int32_t tot;
    int32_t at = wg_exclusive_scan<PT>(sum, s.scan, &tot);
    for (int32_t r = r0; r < r1; r++) {
        const int32_t k = s.evStart[r];
        s.evStart[r] = (uint16_t)at;
        at += k;
    }
    __syncthreads();
    // every event to the next free slot of its run, from the bucket
    for (int32_t g = tid; g < n; g += PT) {
        const uint64_t key = a.bkKey[base + g];
        const float add = a.bkAdd[base + g];
        const uint32_t q = run_inc(s, run_of_key(s, key));
        s.evKey[q] = key;
        s.evAdd[q] = add;
    }
    __syncthreads();
    if (n > GREG * PT) {
        sort_runs(a, s, nRuns, base);
        return n;
    }
    // each run in key order: thread t ranks the events at positions t, t+PT,
    // ... within their runs (the lanes of a wave mostly share a run, so the
    // keys they compare are LDS broadcasts), then, once every rank is known,
    // moves them
    uint64_t pk[GREG];
    float pa[GREG];
    int32_t pd[GREG];
#pragma unroll
    for (int v = 0; v < GREG; v++) {
        const int32_t q = tid + v * PT;
        pd[v] = -1;
        pk[v] = 0;
        pa[v] = 0.f;
        if (q < n) {
            const uint64_t k = s.evKey[q];
            const int32_t r = run_of_key(s, k);
            const int32_t b0 = r ? s.evStart[r - 1] : 0, b1 = s.evStart[r];
            if (b1 - b0 > 1) {
                int32_t rank = 0;
                for (int32_t h = b0; h < b1; h++) rank += s.evKey[h] < k;   // keys are distinct
                if (b0 + rank != q) { pd[v] = b0 + rank; pk[v] = k; pa[v] = s.evAdd[q]; }
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int v = 0; v < GREG; v++)
        if (pd[v] >= 0) { s.evKey[pd[v]] = pk[v]; s.evAdd[pd[v]] = pa[v]; }
    __syncthreads();
    return n;
}

// ------------------------------------------------------------------ cells

struct Cell {
    NeuronState st;
    float bias, thr, sign;
    uint32_t count;   // spike count increments this launch
    uint32_t mask;    // steps of this epoch at which the neuron spiked
    uint32_t prev;    // the same for the previous epoch: its arrivals now
    uint32_t forced;  // driven cells: steps whose spike was forced
    int32_t i;        // neuron index, -1 if the slot is empty
};

// No branch: an empty slot (i = -1) reads neuron 0 and is never used or
// stored, so every cell's loads can be in flight at once.
__device__ __forceinline__ void load_cell(const DevBrain &d, Cell &c, int32_t i) {
    c.i = i;
    c.count = 0; c.mask = 0; c.prev = 0; c.forced = 0;
    const int32_t j = i < 0 ? 0 : i;
    c.st.v = d.v[j]; c.st.gE = d.gE[j]; c.st.gI = d.gI[j]; c.st.refr = d.refr[j];
    c.st.trace = d.trace[j]; c.st.adapt = d.adapt[j]; c.st.res = d.res[j];
    c.bias = d.bias[j]; c.thr = d.thr[j]; c.sign = d.g.sign[j];
}

__device__ __forceinline__ void store_cell(const DevBrain &d, Cell &c) {
    if (c.i < 0) return;
    const int32_t i = c.i;
    d.v[i] = c.st.v; d.gE[i] = c.st.gE; d.gI[i] = c.st.gI; d.refr[i] = c.st.refr;
    d.trace[i] = c.st.trace; d.adapt[i] = c.st.adapt; d.res[i] = c.st.res;
    d.spikeCount[i] += c.count;
    c.count = 0;
}

// lif.c phase 1 for one neuron: depression if it arrives now, then this
// step's inputs (`has`: the neuron's bit of hasEv at step m)
__device__ __forceinline__ void cell_inputs_has(const Consts &k, Cell &c, const Shared &s, int32_t nev, int32_t m,
                                                uint32_t ql, bool has) {
    if ((c.prev >> m) & 1u) {
        if (c.sign * c.st.res != 0.f) c.st.res = c.st.res - k.depU * c.st.res;
    }
    if (has) apply_run(s, (uint32_t)m, ql, c.st.gE, c.st.gI);
}
__device__ __forceinline__ void cell_inputs(const Consts &k, Cell &c, const Shared &s, int32_t nev, int32_t m,
                                            uint32_t ql) {
    cell_inputs_has(k, c, s, nev, m, ql, (s.hasEv[m * WPS + (ql >> 5)] >> (ql & 31)) & 1u);
}

// A driven block's final cell state -> the stage (coalesced) at the launch's
// end, where the next lookahead launch finds it (k_unstage writes it to the
// brain's arrays when anything else needs them).
__device__ __forceinline__ void stage_out(const PersistArgs &a, const Cell (&cells)[CELLS], int32_t qb, int32_t P) {
    float *st = a.dstage;
    const int tid = threadIdx.x;
#pragma unroll
    for (int u = 0; u < DPT; u++) {
        const Cell &c = cells[u];
        if (c.i < 0) continue;
        const int32_t q = qb + tid * P + u;
        st[0 * DRV_CAP + q] = c.st.v; st[1 * DRV_CAP + q] = c.st.gE; st[2 * DRV_CAP + q] = c.st.gI;
        st[3 * DRV_CAP + q] = c.st.refr; st[4 * DRV_CAP + q] = c.st.trace; st[5 * DRV_CAP + q] = c.st.adapt;
        st[6 * DRV_CAP + q] = c.st.res;
        st[12 * DRV_CAP + q] = __uint_as_float(c.count);
    }
}

// ------------------------------------------------------------------ simulate

__device__ __forceinline__ void sim_regular(const PersistArgs &a, Shared &s, Cell (&cells)[CELLS], int32_t nev, int32_t par,
                            int32_t L, int32_t lo) {
    const DevBrain &d = a.d;
    const int lane = threadIdx.x & 63;
    uint64_t *spk = a.spk + (int64_t)par * D_MAX * a.nwords;
    for (int32_t m = 0; m < L; m++) {
#pragma unroll
        for (int u = 0; u < NPT; u++) {
            Cell &c = cells[u];
            bool fire = false;
            if (c.i >= 0) {
                cell_inputs(d.k, c, s, nev, m, (uint32_t)(c.i - lo));
                membrane(d.k, c.st, c.bias);
                if (crosses(d.k, c.st, c.thr)) {
                    spike_effects(d.k, c.st, d.k.tRef);
                    c.count++;
                    c.mask |= 1u << m;
                    fire = true;
                }
            }
            // the wave's 64 neurons are consecutive and 64-aligned, and no
            // other block writes this word: a plain store
            const uint64_t bits = __ballot(fire);
            if (lane == 0 && bits) {
                const int32_t first = lo + (int32_t)(threadIdx.x & ~63) + u * PT;
                spk[(int64_t)m * a.nwords + (first >> 6)] = bits;
            }
        }
    }
}

__device__ __forceinline__ void sim_driven(const PersistArgs &a, Shared &s, Cell (&cells)[CELLS], const float (&pFire)[DPT],
                           int32_t nev, int32_t par, int32_t L, uint32_t &rng) {
    const DevBrain &d = a.d;
    const int tid = threadIdx.x;
    const int32_t nD = d.nDriven;
    const int32_t P = (nD + PT - 1) / PT;
    const float forcedRefr = d.k.tRef + d.k.dt;
    const bool timed = tid == 0 && (a.flags & 1);
#define FB_SUB(k)                                      \
    if (timed) {                                       \
        const uint64_t now = wall_clock64();           \
        s.sub[k] += now - s.sub[6];                    \
        s.sub[6] = now;                                \
    }
    if (timed) s.sub[6] = wall_clock64();
    int32_t nPrev = 0;
    for (int32_t m = 0; m < L; m++) {
        // lif.c phase 1: arrivals for the driven neurons.  The thread's cells
        // are the consecutive local targets q0 .. q0+P-1 (P <= DPT): their
        // hasEv bits come from one word, or two when they straddle one.
        const uint32_t q0 = (uint32_t)(tid * P);
        const int32_t w0 = m * WPS + (int32_t)(q0 >> 5);
        uint64_t pair = s.hasEv[w0];
        if ((q0 & 31u) + (uint32_t)P > 32u) pair |= (uint64_t)s.hasEv[w0 + 1] << 32;
        const uint32_t evBits = (uint32_t)(pair >> (q0 & 31u));
        uint32_t nr = 0;   // bit u: cell u is not refractory and draws
#pragma unroll
        for (int u = 0; u < DPT; u++) {
            Cell &c = cells[u];
            if (c.i >= 0) {
                cell_inputs_has(d.k, c, s, nev, m, q0 + (uint32_t)u, (evBits >> u) & 1u);
                if (!(c.st.refr > 0)) nr |= 1u << u;
            }
        }
        FB_SUB(0)
        // lif.c phase 2: one draw per non-refractory driven neuron, in list order
        int32_t n;
        const int32_t off = wg_scan_flags<PT, DPT>(nr, s.scan + (m & 1) * 16, &n);
        if (nPrev > 0) rng = s.rngNext[(m - 1) & 1];   // the last draw of the previous step
        nPrev = n;
        FB_SUB(1)
        FB_SUB(2)
        int32_t j = off;
        // this thread's first draw is at stream position off+1 after rng:
        // M^(off%32+1) (M^(32 (off/32)) rng), off < 2048; the later ones are single steps
        uint32_t x = nr ? matvec(s.leap + (off & 31) * LEAP_STRIDE, matvec(s.leap + (32 + (off >> 5)) * LEAP_STRIDE, rng))
                        : 0u;
        bool first = true;
#pragma unroll
        for (int u = 0; u < DPT; u++) {
            Cell &c = cells[u];
            if ((nr >> u) & 1u) {
                if (!first) x = xs32(x);
                first = false;
                if (j == n - 1) s.rngNext[m & 1] = x;
                j++;
                const float uu = (float)(x >> 8) * (1.0f / 16777216.0f);
                if (uu < pFire[u]) {
                    c.st.v = d.k.vReset;
                    c.st.refr = forcedRefr;
                    c.st.trace = 1.f;
                    c.st.adapt = c.st.adapt + d.k.adaptInc;
                    c.count++;
                    c.mask |= 1u << m;
                    c.forced |= 1u << m;
                }
            }
        }
        FB_SUB(3)
        // lif.c phases 3-4 for the driven neurons
#pragma unroll
        for (int u = 0; u < DPT; u++) {
            Cell &c = cells[u];
            if (c.i >= 0) {
                membrane(d.k, c.st, c.bias);
                if (crosses(d.k, c.st, c.thr)) {
                    spike_effects(d.k, c.st, d.k.tRef);
                    c.count++;
                    c.mask |= 1u << m;
                }
            }
        }
        FB_SUB(4)
    }
    __syncthreads();
    if (nPrev > 0) rng = s.rngNext[(L - 1) & 1];
    // the epoch's spike steps of every driven neuron, for the list builders
#pragma unroll
    for (int u = 0; u < DPT; u++) {
        const Cell &c = cells[u];
        if (c.i >= 0) {
            a.dmask[par * DRV_CAP + tid * P + u] = c.forced;
            a.tmask[par * DRV_CAP + tid * P + u] = c.mask & ~c.forced;
            a.dres[par * DRV_CAP + tid * P + u] = c.st.res;
        }
    }
    FB_SUB(5)
#undef FB_SUB
}

// ------------------------------------------------------------------ several driven blocks

// Driven block `me` of G: its count of non-refractory neurons for launch step
// t goes to xc[me][t % XC_RING] as (t + 1) << 32 | count.
__device__ __forceinline__ void xc_publish(const PersistArgs &a, int32_t me, int32_t t, int32_t n) {
    uint64_t *w = a.xc + (int64_t)me * XC_RING + (t % XC_RING);
    (void)__hip_atomic_exchange(w, ((uint64_t)(uint32_t)(t + 1) << 32) | (uint32_t)n, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
}

// Waits for driven block k's count of step t (every lane of the wave loads
// the same word).  A block publishes step t's count before it can need any
// count of step t, and no block runs more than two steps ahead of another,
// so the ring slot holds step t's word or an older one.  The wait is bounded
// like the grid barrier's: on a timeout it sets the flag, gives 0 and stops
// waiting for the rest of the launch (`dead`); the next epoch boundary stops
// every block.
__device__ __forceinline__ uint32_t xc_wait(const PersistArgs &a, int32_t k, int32_t t, bool &dead) {
    const uint64_t *w = a.xc + (int64_t)k * XC_RING + (t % XC_RING);
    const uint32_t tag = (uint32_t)(t + 1);
    uint64_t v = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((uint32_t)(v >> 32) == tag || dead) return dead ? 0u : (uint32_t)v;
    const uint64_t t0 = wall_clock64();
    while ((uint32_t)(v >> 32) != tag) {
        __builtin_amdgcn_s_sleep(1);
        if (wall_clock64() - t0 > GBAR_TIMEOUT) {
            atomicOr(&a.ctl[1], GBAR_TIMED_OUT);
            dead = true;
            return 0u;
        }
        v = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return (uint32_t)v;
}

// This block's non-refractory flags (bit u: cell u draws) and its ballot
// scan; every thread of the block calls it.
__device__ __forceinline__ uint32_t driven_flags(const Cell (&cells)[CELLS]) {
    uint32_t nr = 0;
#pragma unroll
    for (int u = 0; u < DPT; u++)
        if (cells[u].i >= 0 && !(cells[u].st.refr > 0)) nr |= 1u << u;
    return nr;
}

// The driven block of several: as sim_driven, with the stream position of a
// draw = (counts of the blocks before this one) + (position in this block),
// and the stream state of step t = M^(total draws of step t-1) applied to
// that of step t-1.  `t0`: the launch step of the epoch's first step; `nr`,
// `off`: this block's flags and offsets for step t0, computed and published
// at the end of the previous step (or before the first epoch), and left for
// the step after the epoch.
__device__ __forceinline__ void sim_driven_x(const PersistArgs &a, Shared &s, Cell (&cells)[CELLS], const float (&pFire)[DPT],
                             int32_t nev, int32_t par, int32_t L, int32_t t0, uint32_t &rng, uint32_t &nr,
                             int32_t &off, bool &dead) {
    const DevBrain &d = a.d;
    const int tid = threadIdx.x;
    const int32_t me = blockIdx.x;
    const int32_t qb = me * a.Q;
    const int32_t nDk = min(a.Q, d.nDriven - qb);
    const int32_t P = (nDk + PT - 1) / PT;
    const float forcedRefr = d.k.tRef + d.k.dt;
    const bool timed = tid == 0 && (a.flags & 1);
#define FB_SUB(k)                                      \
    if (timed) {                                       \
        const uint64_t now = wall_clock64();           \
        s.sub[k] += now - s.sub[6];                    \
        s.sub[6] = now;                                \
    }
    if (timed) s.sub[6] = wall_clock64();
    for (int32_t m = 0; m < L; m++) {
        const int32_t t = t0 + m;
        // lif.c phase 1: arrivals for the driven neurons (local targets
        // q0 .. q0+P-1 of this block)
        const uint32_t q0 = (uint32_t)(tid * P);
        const int32_t w0 = m * WPS + (int32_t)(q0 >> 5);
        uint64_t pair = s.hasEv[w0];
        if ((q0 & 31u) + (uint32_t)P > 32u) pair |= (uint64_t)s.hasEv[w0 + 1] << 32;
        const uint32_t evBits = (uint32_t)(pair >> (q0 & 31u));
#pragma unroll
        for (int u = 0; u < DPT; u++) {
            Cell &c = cells[u];
            if (c.i >= 0) cell_inputs_has(d.k, c, s, nev, m, q0 + (uint32_t)u, (evBits >> u) & 1u);
        }
        FB_SUB(0)
        // the other blocks' counts: this step's before this block, and the
        // previous step's total for the stream state
        uint32_t before = 0, prevTotal = 0;
        for (int32_t k = 0; k < a.G; k++) {
            if (k < me) before += xc_wait(a, k, t, dead);
            if (t > 0) prevTotal += xc_wait(a, k, t - 1, dead);
        }
        if (prevTotal) rng = leap_jump(s.leap, prevTotal, rng);
        FB_SUB(1)
        FB_SUB(2)
        // lif.c phase 2: one draw per non-refractory driven neuron, in list order
        uint32_t x = nr ? leap_jump(s.leap, before + (uint32_t)off + 1u, rng) : 0u;
        bool first = true;
#pragma unroll
        for (int u = 0; u < DPT; u++) {
            Cell &c = cells[u];
            if ((nr >> u) & 1u) {
                if (!first) x = xs32(x);
                first = false;
                const float uu = (float)(x >> 8) * (1.0f / 16777216.0f);
                if (uu < pFire[u]) {
                    c.st.v = d.k.vReset;
                    c.st.refr = forcedRefr;
                    c.st.trace = 1.f;
                    c.st.adapt = c.st.adapt + d.k.adaptInc;
                    c.count++;
                    c.mask |= 1u << m;
                    c.forced |= 1u << m;
                }
            }
        }
        FB_SUB(3)
        // lif.c phases 3-4 for the driven neurons
#pragma unroll
        for (int u = 0; u < DPT; u++) {
            Cell &c = cells[u];
            if (c.i >= 0) {
                membrane(d.k, c.st, c.bias);
                if (crosses(d.k, c.st, c.thr)) {
                    spike_effects(d.k, c.st, d.k.tRef);
                    c.count++;
                    c.mask |= 1u << m;
                }
            }
        }
        // the next step's flags, fixed now: publish this block's count
        nr = driven_flags(cells);
        int32_t n;
        off = wg_scan_flags<PT, DPT>(nr, s.scan + (m & 1) * 16, &n);
        if (tid == 0) xc_publish(a, me, t + 1, n);
        FB_SUB(4)
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < DPT; u++) {
        const Cell &c = cells[u];
        if (c.i >= 0) {
            const int32_t q = qb + tid * P + u;
            a.dmask[par * DRV_CAP + q] = c.forced;
            a.tmask[par * DRV_CAP + q] = c.mask & ~c.forced;
            a.dres[par * DRV_CAP + q] = c.st.res;
        }
    }
    FB_SUB(5)
#undef FB_SUB
}

// ------------------------------------------------------------------ push: own spikes -> next epoch's buckets

// A regular block's spikes of the epoch -> the LDS push list (s.spkI /
// spkK / spkR); returns the row count (at most SPK_CAP, flag 4 beyond).
__device__ __forceinline__ int32_t collect_spikes(const PersistArgs &a, Shared &s, Cell (&cells)[CELLS], int32_t par) {
    const DevBrain &d = a.d;
    const int tid = threadIdx.x;
    if (tid == 0) s.nspk = 0;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < CELLS; u++) {
        const Cell &c = cells[u];
        uint32_t bits = c.mask;
        while (bits) {
            const int m = __ffs(bits) - 1;
            bits &= bits - 1;
            const int32_t e = atomicAdd(&s.nspk, 1);
            if (e < SPK_CAP) {
                s.spkI[e] = c.i;
                s.spkK[e] = ((uint32_t)m << 24) | (KEY_THRESHOLD + (uint32_t)c.i);
                // sign x res at its arrival step of the next epoch
                s.spkR[e] = c.sign * res_after(d.k, c.st.res, m);
            }
        }
    }
    __syncthreads();
    const int32_t ns = s.nspk;
    if (tid == 0 && ns > SPK_CAP) flag_overflow(a, par, 4);
    return min(ns, SPK_CAP);
}

// A regular block pushes its own spikes.
// A regular block pushes its own spikes (buckets of parity par); rows longer
// than LONG_ROW go to the shared list of parity lpar.
__device__ __forceinline__ void push_spikes(const PersistArgs &a, Shared &s, Cell (&cells)[CELLS], int32_t par, int32_t lpar) {
    const int32_t ns = collect_spikes(a, s, cells, par);
    push_rows(a, s, par, ns, lpar);
    publish_stats(a, s);
}

// The driven block's rows are many (every forced spike), so the SHARE_BLOCKS
// blocks after the driven ones push them (share_index), each for a fixed
// range of driven-list positions whose rows it holds in LDS (share_setup).
// Not the driven blocks themselves: their serial chain is the launch's
// critical path, from the prologue on.  Once the driven blocks
// have published the epoch (PersistArgs::epc), a share block reads the spike
// steps of its neurons (dmask: forced, tmask: threshold, as the driven block
// wrote them) and their res, and writes their rows' events.
// This block's index among the share blocks, or -1.
__device__ __forceinline__ int32_t share_index(const PersistArgs &a) {
    const int32_t b = (int32_t)blockIdx.x - a.G;
    return b >= 0 && b < min((int32_t)gridDim.x - a.G, SHARE_BLOCKS) ? b : -1;
}

__device__ __forceinline__ void share_setup(const PersistArgs &a, Shared &s) {
    const DevBrain &d = a.d;
    const int tid = threadIdx.x;
    const int32_t nsh = min((int32_t)gridDim.x - a.G, SHARE_BLOCKS), b = share_index(a);
    if (b < 0) return;
    const int32_t q0 = (int32_t)((int64_t)d.nDriven * b / nsh), q1 = (int32_t)((int64_t)d.nDriven * (b + 1) / nsh);
    if (tid == 0) { s.shQ0 = q0; s.shNq = q1 - q0; }
    for (int32_t k = tid; k < q1 - q0; k += PT) {
        const int32_t i = d.driven[q0 + k];
        s.shI[k] = i;
        s.shRA[k] = d.g.indptr[i];
        s.shLen[k] = (int32_t)(d.g.indptr[i + 1] - d.g.indptr[i]);
        s.shSyn[k] = d.g.rowSyn[i];
        s.shSign[k] = d.g.sign[i];
    }
}

// `par`: the parity of the buckets written (the next epoch's); `mpar`: the
// parity of the driven masks read (this epoch's).
__device__ __forceinline__ void push_driven_share(const PersistArgs &a, Shared &s, int32_t par, int32_t mpar) {
    const DevBrain &d = a.d;
    const int tid = threadIdx.x;
    const bool timed = tid == 0 && (a.flags & 1);
#define FB_SHARE(k)                                    \
    if (timed) {                                       \
        const uint64_t now = wall_clock64();           \
        s.share[k] += now - s.share[3];                \
        s.share[3] = now;                              \
    }
    if (timed) s.share[3] = wall_clock64();
    if (share_index(a) < 0) return;   // uniform in the block
    __syncthreads();        // the LDS push list may still be read by this block's own push
    if (tid == 0) s.nspk = 0;
    __syncthreads();
    const int32_t q0 = s.shQ0, nq = s.shNq;
    // rows: one per spike step of each neuron (a forced and a threshold spike never share a step)
    for (int32_t k = tid; k < nq; k += PT) {
        const uint32_t fm = a.dmask[mpar * DRV_CAP + q0 + k];
        uint32_t bits = fm | a.tmask[mpar * DRV_CAP + q0 + k];
        if (!bits) continue;
        const float res = a.dres[mpar * DRV_CAP + q0 + k];
        const int32_t i = s.shI[k];
        while (bits) {
            const int m = __ffs(bits) - 1;
            bits &= bits - 1;
            const int32_t e = atomicAdd(&s.nspk, 1);   // beyond SPK_CAP rows: flag 4 below
            if (e < SPK_CAP) {
                s.spkI[e] = k;
                s.spkK[e] = ((uint32_t)m << 24) | (((fm >> m) & 1u) ? (uint32_t)(q0 + k) : KEY_THRESHOLD + (uint32_t)i);
                // sign x res at its arrival step of the next epoch
                s.spkR[e] = s.shSign[k] * res_after(d.k, res, m);
            }
        }
    }
    __syncthreads();
    const int32_t nsAll = s.nspk;
    if (tid == 0 && nsAll > SPK_CAP) flag_overflow(a, par, 4);
    const int32_t ns = min(nsAll, SPK_CAP);
    // row lengths from the LDS table -> the flattened edge list
    int32_t len[RPT];
    int32_t mySum = 0;
    unsigned long long syn = 0;
#pragma unroll
    for (int v = 0; v < RPT; v++) {
        const int32_t r = RPT * tid + v;
        len[v] = 0;
        if (r < ns && s.spkR[r] != 0.f) {
            const int32_t k = s.spkI[r];
            len[v] = s.shLen[k];
            syn += s.shSyn[k];
        }
        mySum += len[v];
    }
    int32_t total;
    int32_t off = wg_exclusive_scan<PT>(mySum, s.scan, &total);
#pragma unroll
    for (int v = 0; v < RPT; v++) {
        const int32_t r = RPT * tid + v;
        if (r < ns) { s.rowOff[r] = off; s.rowA[r] = s.shRA[s.spkI[r]]; }
        off += len[v];
    }
    if (tid == 0) s.rowOff[ns] = total;
    if (mySum) {
        atomicAdd(&d.sc->edgeEvents, (unsigned long long)mySum);
        atomicAdd(&d.sc->synEvents, syn);
    }
    __syncthreads();
    FB_SHARE(0)
    push_edges(a, s, par, ns, 0, total);
    __syncthreads();
    FB_SHARE(1)
    FB_SHARE(2)
#undef FB_SHARE
}

// ------------------------------------------------------------------ lists: one step's spikes in lif.c order

__device__ __forceinline__ void build_list(const PersistArgs &a, Shared &s, int32_t par, int32_t t, int32_t m) {
    const DevBrain &d = a.d;
    const int tid = threadIdx.x;
    const int32_t nD = d.nDriven;
    const int32_t slot = (head_at(a, t) + d.nslots - 1) % d.nslots;
    int32_t *out = d.ring + (int64_t)slot * d.ringStride;
    const uint32_t *fl = a.dmask + par * DRV_CAP;
    const uint32_t *tf = a.tmask + par * DRV_CAP;
    uint64_t *bits = a.spk + ((int64_t)par * D_MAX + m) * a.nwords;
    const int32_t P = (nD + PT - 1) / PT;
    const int32_t dlo = min(nD, tid * P), dhi = min(nD, dlo + P);
    // Each loop issues a batch of loads before it uses any: a loop that used
    // each load before issuing the next paid ~1 us per word at a launch's end.
    constexpr int BQ = 4, BW = 2;
    int32_t nf = 0;
    for (int32_t q0 = dlo; q0 < dhi; q0 += BQ) {
        uint32_t f[BQ], th[BQ];
        int32_t ii[BQ];
#pragma unroll
        for (int j = 0; j < BQ; j++) {
            const int32_t q = min(q0 + j, dhi - 1);
            f[j] = fl[q]; th[j] = tf[q]; ii[j] = d.driven[q];
        }
#pragma unroll
        for (int j = 0; j < BQ; j++) {
            if (q0 + j >= dhi) break;
            nf += (f[j] >> m) & 1u;
            if ((th[j] >> m) & 1u)   // threshold spikes of driven neurons join the bitmask
                atomicOr((unsigned long long *)&bits[ii[j] >> 6], 1ull << (ii[j] & 63));
        }
    }
    // the bitmask's atomics are complete before any thread reads it (a fence
    // with only them in flight: the list's own stores come after it)
    __threadfence_block();
    int32_t ndTot;
    int32_t fpos = wg_exclusive_scan<PT>(nf, s.scan, &ndTot);
    const int32_t W = (a.nwords + PT - 1) / PT;
    const int32_t wlo = min(a.nwords, tid * W), whi = min(a.nwords, wlo + W);
    // the words are kept in LDS for the second pass when they fit (the
    // events' space is free between an epoch's push and the next gather)
    uint64_t *keep = a.nwords <= EV_CAP ? s.evKey : nullptr;
    int32_t cnt = 0;
    for (int32_t w0 = wlo; w0 < whi; w0 += BW) {
        uint64_t v[BW];
#pragma unroll
        for (int j = 0; j < BW; j++)
            v[j] = w0 + j < whi ? __hip_atomic_load(&bits[w0 + j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0ull;
#pragma unroll
        for (int j = 0; j < BW; j++) {
            cnt += __popcll(v[j]);
            if (keep && w0 + j < whi) keep[w0 + j] = v[j];
        }
    }
    // the forced spikes first in the list, in driven-list order
    for (int32_t q0 = dlo; q0 < dhi; q0 += BQ) {
        uint32_t f[BQ];
        int32_t ii[BQ];
#pragma unroll
        for (int j = 0; j < BQ; j++) {
            const int32_t q = min(q0 + j, dhi - 1);
            f[j] = fl[q]; ii[j] = d.driven[q];
        }
#pragma unroll
        for (int j = 0; j < BQ; j++)
            if (q0 + j < dhi && ((f[j] >> m) & 1u)) out[fpos++] = ii[j];
    }
    int32_t thTot;
    int32_t pos = ndTot + wg_exclusive_scan<PT>(cnt, s.scan, &thTot);
    for (int32_t w0 = wlo; w0 < whi; w0 += BW) {
        uint64_t v[BW];
#pragma unroll
        for (int j = 0; j < BW; j++)
            v[j] = w0 + j >= whi ? 0ull
                   : keep        ? keep[w0 + j]
                                 : __hip_atomic_load(&bits[w0 + j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
        for (int j = 0; j < BW; j++) {
            uint64_t b = v[j];
            if (!b) continue;
            const int32_t w = w0 + j;
            bits[w] = 0ull;
            while (b) {
                const int bit = __ffsll((unsigned long long)b) - 1;
                out[pos++] = (w << 6) + bit;
                b &= b - 1;
            }
        }
    }
    if (tid == 0) {
        d.ringCount[slot] = ndTot + thTot;
        atomicAdd(&d.sc->spikes, (unsigned long long)(ndTot + thTot));
    }
    __syncthreads();
}

// lists of epoch steps [t0, t0+L), spread over the spare blocks (or the
// regular blocks when the grid has no spare ones)
__device__ __forceinline__ void build_lists(const PersistArgs &a, Shared &s, int32_t par, int32_t t0, int32_t L) {
    const int32_t first = a.nSpare > 0 ? a.S : a.G;
    const int32_t count = a.nSpare > 0 ? a.nSpare : max(1, a.S - a.G);
    const int32_t b = (int32_t)blockIdx.x - first;
    if (b < 0 || b >= count) return;
    for (int32_t m = b; m < L; m += count) build_list(a, s, par, t0 + m, m);
    if (b == 0 && threadIdx.x == 0) {
        // the slot delivered at the epoch's last step is empty until a later list fills it
        a.d.ringCount[head_at(a, t0 + L - 1)] = 0;
    }
}

// ------------------------------------------------------------------ the driven neurons' state

// Fields of the driven stage, in driven-list order: the cell state (0-6),
// then bias, threshold, sign and drive, then the neuron index (as bits), then
// the spike-count increments the brain's spikeCount does not hold yet (as
// bits).  At a launch's end the driven blocks leave their cells' state and
// increments there, and the stage stays the driven neurons' state between
// lookahead launches: the brain's arrays are brought up to date (k_unstage)
// only when something else reads or writes them.  So a run of lookahead
// calls pays no hand-over at either end: ~1,650 scattered neurons' eight
// fields cost a launch's end 10-15 us even spread over the regular blocks,
// each dependent global load there taking 3-5 us.
constexpr int DSTAGE_F = 13;

// Every block but the driven ones copies a share of the driven neurons' state
// and parameters into the stage, the parts the flags do not mark as there
// already (block b of nb: positions b*PT + tid, b*PT + tid + nb*PT, ...).
// Read by one workgroup, those scattered lines take ~20 us at a launch's
// start; spread over the grid they are one round trip, and the driven blocks
// read the stage coalesced after the barrier (before it, when the stage is
// whole already).
__device__ __forceinline__ void stage_driven(const PersistArgs &a, int32_t b, int32_t nb) {
    const DevBrain &d = a.d;
    float *st = a.dstage;
    const bool state = !(a.flags & STAGE_STATE), params = !(a.flags & STAGE_PARAMS);
    if (!state && !params) return;
    for (int32_t q = b * PT + (int32_t)threadIdx.x; q < d.nDriven; q += nb * PT) {
        const int32_t i = d.driven[q];
        if (state) {
            st[0 * DRV_CAP + q] = d.v[i];
            st[1 * DRV_CAP + q] = d.gE[i];
            st[2 * DRV_CAP + q] = d.gI[i];
            st[3 * DRV_CAP + q] = d.refr[i];
            st[4 * DRV_CAP + q] = d.trace[i];
            st[5 * DRV_CAP + q] = d.adapt[i];
            st[6 * DRV_CAP + q] = d.res[i];
            st[12 * DRV_CAP + q] = 0.f;
        }
        if (params) {
            st[7 * DRV_CAP + q] = d.bias[i];
            st[8 * DRV_CAP + q] = d.thr[i];
            st[9 * DRV_CAP + q] = d.g.sign[i];
            st[10 * DRV_CAP + q] = d.drive[i];
            st[11 * DRV_CAP + q] = __int_as_float(i);
        }
    }
}

// A driven cell from the stage (q = driven-list position, -1 for an empty slot).
__device__ __forceinline__ void load_staged_cell(const PersistArgs &a, Cell &c, float &pFire, int32_t q) {
    const float *st = a.dstage;
    const int32_t j = q < 0 ? 0 : q;
    c.mask = 0; c.prev = 0; c.forced = 0;
    c.count = __float_as_uint(st[12 * DRV_CAP + j]);   // increments the brain's spikeCount lacks
    c.st.v = st[0 * DRV_CAP + j]; c.st.gE = st[1 * DRV_CAP + j]; c.st.gI = st[2 * DRV_CAP + j];
    c.st.refr = st[3 * DRV_CAP + j]; c.st.trace = st[4 * DRV_CAP + j]; c.st.adapt = st[5 * DRV_CAP + j];
    c.st.res = st[6 * DRV_CAP + j];
    c.bias = st[7 * DRV_CAP + j]; c.thr = st[8 * DRV_CAP + j]; c.sign = st[9 * DRV_CAP + j];
    const float dr = st[10 * DRV_CAP + j];
    c.i = q < 0 ? -1 : __float_as_int(st[11 * DRV_CAP + j]);
    pFire = q < 0 ? 0.f : dr * a.d.k.dtS;
}

// ------------------------------------------------------------------ the kernel

__global__ void __launch_bounds__(PT) k_persist(PersistArgs a) {
    __shared__ Shared s;
    const DevBrain &d = a.d;
    const int tid = threadIdx.x;
    const int32_t me = blockIdx.x;
    const bool drivenBlock = me < a.G;
    const bool owning = me < a.S;
    const int32_t lo = drivenBlock || !owning ? 0 : a.bstart[me - a.G];
    const int32_t hi = drivenBlock || !owning ? 0 : a.bstart[me - a.G + 1];

    // clocks of the first driven block, the first regular block and the first spare block
    const int32_t timedSlot = me == 0 ? 0 : me == a.G ? 1 : (a.nSpare > 0 && me == a.S) ? 2 : -1;
    const bool timed = tid == 0 && timedSlot >= 0;
    const uint64_t tStart = wall_clock64();
    if (timed) {
        for (int k = 0; k < NPH; k++) s.ph[k] = 0;
        for (int k = 0; k < 6; k++) s.sub[k] = 0;
        s.ph[NPH] = wall_clock64();
    }
    if (tid == 0) s.share[0] = s.share[1] = s.share[2] = 0;
#define FB_PHASE(k)                                    \
    if (timed) {                                       \
        const uint64_t now = wall_clock64();           \
        s.ph[k] += now - s.ph[NPH];                    \
        s.ph[NPH] = now;                               \
    }

    if (tid == 0) { s.ev = 0; s.syn = 0; s.rngNext[0] = s.rngNext[1] = 0; }
    Cell cells[CELLS];
    float pFire[DPT];
    uint32_t rng = 0;
    uint32_t xNr = 0;      // several driven blocks: the next step's flags and offset (sim_driven_x)
    int32_t xOff = 0;
    bool xDead = false;
    // the driven cells come from the stage: at once when the last launch left
    // it whole, else after the barrier that ends the other blocks' staging
    const bool stagedEarly = (a.flags & (STAGE_STATE | STAGE_PARAMS)) == (STAGE_STATE | STAGE_PARAMS);
    auto loadDriven = [&]() {
        const int32_t qb = me * a.Q;
        const int32_t nDk = min(a.Q, d.nDriven - qb);
        const int32_t P = (nDk + PT - 1) / PT;
#pragma unroll
        for (int u = 0; u < DPT; u++) {
            const int32_t q = tid * P + u;
            load_staged_cell(a, cells[u], pFire[u], u < P && q < nDk ? qb + q : -1);
        }
    };
    if (drivenBlock) {
        leap_load<PT, LEAP_PERSIST_MATS>(s.leap, d.leap);
        rng = d.sc->rng;
        if (stagedEarly) loadDriven();
    } else {
#pragma unroll
        for (int u = 0; u < CELLS; u++) {
            const int32_t i = lo + tid + u * PT;
            load_cell(d, cells[u], (u < NPT && owning && i < hi && a.drvPos[i] < 0) ? i : -1);
        }
        stage_driven(a, me - a.G, (int32_t)gridDim.x - a.G);
    }
    share_setup(a, s);
    __syncthreads();

    const int32_t L0 = min(a.D, a.nsteps);
    if (owning && !drivenBlock) prologue_scatter(a, s, L0);
    FB_PHASE(5)
    if (!grid_barrier(a.bar, &a.ctl[1])) return;
    FB_PHASE(4)
    if (drivenBlock) {
        if (!stagedEarly) loadDriven();
        if (a.G > 1) {
            // the first step's count
            xNr = driven_flags(cells);
            int32_t n;
            xOff = wg_scan_flags<PT, DPT>(xNr, s.scan + 32, &n);
            if (tid == 0) xc_publish(a, me, 0, n);
        }
    }
    if (d.k.depU != 0.f && owning) {
#pragma unroll
        for (int u = 0; u < CELLS; u++)
            if (cells[u].i >= 0) { cells[u].prev = a.arrMask[cells[u].i]; a.arrMask[cells[u].i] = 0u; }
    }

    int32_t done = 0, e = 0, prevL = 0;
    bool listsBuilt = false;   // the last epoch's lists were built before a stop
    {
        while (done < a.nsteps) {
            const int32_t L = min(a.D, a.nsteps - done);
            const int32_t par = e & 1;
            // overflow flags of the pushes into this epoch's buckets (the
            // previous epoch's, the prologue's at e = 0): read now, acted on
            // after the gather (which only reads), so the load overlaps the
            // gather's
            const int32_t stop = __hip_atomic_load(&a.pflag[par], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // the long-row list of the other parity was last read before the
            // previous epoch's barrier and is next written in the next epoch:
            // clear it now
            if (me == 0 && tid == 0) (void)atomicExch(&a.lrCnt[par ^ 1], 0);
            if (e > 0) build_lists(a, s, par ^ 1, done - prevL, prevL);
            FB_PHASE(3)
            int32_t nev = 0;
            if (owning) nev = gather(a, s, par, me);
            FB_PHASE(0)
            if (stop) {   // every block read the same flags after the barrier
                listsBuilt = true;
                break;
            }
            if (owning) {
                if (drivenBlock && a.G > 1) sim_driven_x(a, s, cells, pFire, nev, par, L, done, rng, xNr, xOff, xDead);
                else if (drivenBlock) sim_driven(a, s, cells, pFire, nev, par, L, rng);
                else sim_regular(a, s, cells, nev, par, L, lo);
                FB_PHASE(1)
                const bool more = done + L < a.nsteps;
                if (more && !drivenBlock) push_spikes(a, s, cells, par ^ 1, par);
#pragma unroll
                for (int u = 0; u < CELLS; u++) {
                    cells[u].prev = cells[u].mask;
                    cells[u].mask = 0;
                    cells[u].forced = 0;
                }
                if (more) {
                    // a regular workgroup's pushes (and long rows queued) are
                    // done; a driven workgroup's masks of the epoch are written
                    __builtin_amdgcn_s_waitcnt(0);
                    __syncthreads();
                    if (tid == 0) {
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                        if (drivenBlock) {
                            __hip_atomic_store(&a.epc[GBAR_WORDS + me], e + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        } else {
                            // the last of its group's regular workgroups to
                            // arrive counts the group
                            const int32_t g = me / GBAR_GROUP;
                            const int32_t nr = min((g + 1) * GBAR_GROUP, a.S) - max(g * GBAR_GROUP, a.G);
                            const int32_t arr = atomicAdd(&a.epc[64 * (g + 1)], 1);
                            if (arr % nr == nr - 1) atomicAdd(&a.epc[0], 1);
                        }
                    }
                }
                FB_PHASE(2)
            }
            if (done + L < a.nsteps) {
                // the driven rows and the long rows go out before the epoch's
                // one grid barrier: every workgroup waits for all regular
                // workgroups' pushes (the long rows are queued by then), a
                // share workgroup also for the driven workgroups' masks
                if (tid == 0) {
                    const uint64_t t0 = wall_clock64();
                    const bool share = share_index(a) >= 0;
                    // groups that hold regular workgroups (those are G .. S-1)
                    const int32_t ng = a.S > a.G ? (a.S - 1) / GBAR_GROUP - a.G / GBAR_GROUP + 1 : 0;
                    bool late = false;
                    while (!late) {
                        bool ready = __hip_atomic_load(&a.epc[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= (e + 1) * ng;
                        for (int32_t k = 0; share && ready && k < a.G; k++)
                            ready = __hip_atomic_load(&a.epc[GBAR_WORDS + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= e + 1;
                        if (ready) break;
                        __builtin_amdgcn_s_sleep(1);
                        if (wall_clock64() - t0 > GBAR_TIMEOUT) {
                            atomicOr(&a.ctl[1], GBAR_TIMED_OUT);
                            late = true;
                        }
                    }
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                }
                __syncthreads();
                FB_PHASE(7)
                push_driven_share(a, s, par ^ 1, par);
                push_long_share(a, s, par ^ 1, par);
                FB_PHASE(6)
            }
            if (!grid_barrier(a.bar, &a.ctl[1])) return;
            FB_PHASE(4)
            done += L;
            prevL = L;
            e++;
        }
    }
    // the last epoch's lists, and the state back: the regular cells to the
    // brain's arrays, the driven cells to the stage (also after a stop)
    if (e > 0 && !listsBuilt) build_lists(a, s, (e - 1) & 1, done - prevL, prevL);
    FB_PHASE(3)
    if (drivenBlock) {
        stage_out(a, cells, me * a.Q, (min(a.Q, d.nDriven - me * a.Q) + PT - 1) / PT);
    } else if (owning) {
#pragma unroll
        for (int u = 0; u < CELLS; u++) store_cell(d, cells[u]);
    }
    if (owning) {
        if (me == 0) {
            if (a.G > 1 && done > 0) {
                // the stream after the last step: its draws over every driven block
                uint32_t total = 0;
                for (int32_t k = 0; k < a.G; k++) total += xc_wait(a, k, done - 1, xDead);
                if (total) rng = leap_jump(s.leap, total, rng);
            }
            if (tid == 0) d.sc->rng = rng;
        }
    }
#undef FB_PHASE
    if (timed) {
        for (int k = 0; k < NPH; k++)
            atomicAdd((unsigned long long *)&a.prof[timedSlot * NPH + k], (unsigned long long)s.ph[k]);
        if (me == 0)
            for (int k = 0; k < 6; k++)
                atomicAdd((unsigned long long *)&a.prof[PROF_SUB + k], (unsigned long long)s.sub[k]);
    }
    if (tid == 0 && (a.flags & 1)) {
        for (int k = 0; k < 3; k++)
            atomicAdd((unsigned long long *)&a.prof[PROF_SHARE + k], (unsigned long long)s.share[k]);
        atomicMax((unsigned long long *)&a.prof[PROF_SHARE + 3], (unsigned long long)(s.share[0] + s.share[1] + s.share[2]));
    }
    if (me == 0 && tid == 0) a.ctl[0] = done;
    // The last block out hands the results to the host (pinned memory, so
    // the host needs no copy after the launch) and leaves the counters a
    // launch needs at zero -- barrier counters, exchange words, long-row
    // lists, flags, clocks -- so that the next launch needs no clearing
    // kernel.  (After a launch without overflow the bucket counts are zero
    // already: every gather clears the counts it reads, and the last epoch
    // pushes nothing; after an overflow the host clears them.)  A block that
    // left early after a barrier timeout never arrives here, and the host
    // then reads the device words and clears them itself.
    __syncthreads();
    if (tid == 0) {
        unsigned long long *span = (unsigned long long *)&a.prof[PROF_SPAN];
        atomicMax(&span[0], (unsigned long long)~tStart);
        atomicMax(&span[1], (unsigned long long)tStart);
        const unsigned long long tEnd = wall_clock64();
        atomicMax(&span[2], tEnd);
        atomicMax(&span[drivenBlock ? 4 : owning ? 5 : 6], tEnd);
        // the last block out reads ctl[0], which block 0 stored; everything
        // else it reads is atomics, and the state stores need only the
        // kernel's end (an agent-scope release in every block would write
        // back the L2 once per block)
        if (me == 0) __threadfence();
        s.last = atomicAdd(&a.ctl[2], 1) == (int32_t)gridDim.x - 1;
    }
    __syncthreads();
    if (s.last) {
        __threadfence();
        for (int32_t k = tid; k < GBAR_WORDS; k += PT) a.bar[k] = 0;
        if (tid < DRV_BLOCKS * XC_RING) a.xc[tid] = 0;
        for (int32_t k = tid; k < LR_WORDS; k += PT) a.lrCnt[k] = 0;
        if (tid < PERSIST_PROF) {
            unsigned long long *pk = (unsigned long long *)&a.prof[tid];
            a.hostProf[tid] = tid == PROF_SPAN + 3 ? (uint64_t)wall_clock64()
                                                   : __hip_atomic_load(pk, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            a.prof[tid] = 0;
        }
        if (tid == 0) {
            a.hostCtl[0] = __hip_atomic_load(&a.ctl[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            a.hostCtl[1] = __hip_atomic_load(&a.ctl[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            a.ctl[1] = 0;
            a.ctl[2] = 0;
        }
        __syncthreads();
        __threadfence_system();
        if (tid == 0) __hip_atomic_store(&a.hostCtl[2], a.tag, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
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

struct Layout {
    int32_t nb, G, Q, S, nSpare, grid;
    std::vector<int32_t> starts;   // regular block r owns neurons starts[r] .. starts[r+1]-1 (64-aligned; last = N)
};

// Greedy packing of the 64-neuron chunks into consecutive ranges of at most
// CPB chunks whose cost stays within T (a single heavier chunk takes a range
// of its own); the chunk index of each range's start.
static int32_t pack_chunks(const uint64_t *cost, int32_t nC, int32_t CPB, uint64_t T, std::vector<int32_t> *starts) {
    int32_t blocks = 0, inBlk = 0;
    uint64_t cur = 0;
    for (int32_t c = 0; c < nC; c++) {
        if (blocks == 0 || inBlk == CPB || (inBlk > 0 && cur + cost[c] > T)) {
            if (starts) starts->push_back(c);
            blocks++;
            inBlk = 0;
            cur = 0;
        }
        cur += cost[c];
        inBlk++;
    }
    return blocks;
}

// The driven list splits evenly over the fewest driven blocks that hold it.
// Regular blocks own 64-aligned ranges of at most 1024 neurons, as many as
// equal ranges of the largest size that leaves D blocks of the grid spare
// for the list building (or none spare, if it must), with boundaries that
// balance the chunks' cost (DevGraph::chunkCost): a connectome's hubs sit
// together, and equal ranges would hand one block many times the average
// number of events to sort and rows to push.  Cached per graph and grid.
struct LayoutKey {
    const uint64_t *cost;
    int32_t N, G, D, maxB;
    bool operator<(const LayoutKey &o) const {
        return std::tie(cost, N, G, D, maxB) < std::tie(o.cost, o.N, o.G, o.D, o.maxB);
    }
};
static std::map<LayoutKey, std::vector<int32_t>> g_layouts;

static bool layout_for(const DevBrain &d, int32_t D, Layout &l) {
    const int32_t maxB = g_ps.maxBlocks;
    l.G = std::max(1, (d.nDriven + DRV_BLK - 1) / DRV_BLK);
    l.Q = (d.nDriven + l.G - 1) / l.G;
    if (maxB < l.G + 1) return false;
    const int32_t want = std::max(1, maxB - l.G - D);   // regular blocks if D spare blocks fit
    int64_t nb = ((int64_t)(d.N + want - 1) / want + 63) / 64 * 64;
    if (nb > NB_MAX) {
        nb = ((int64_t)(d.N + (maxB - l.G) - 1) / (maxB - l.G) + 63) / 64 * 64;   // no spares
        if (nb > NB_MAX) return false;
    }
    const int32_t need = (int32_t)((d.N + nb - 1) / nb);
    const LayoutKey key{d.g.chunkCost, d.N, l.G, D, maxB};
    auto it = g_layouts.find(key);
    if (it == g_layouts.end()) {
        const int32_t nC = (d.N + 63) / 64, CPB = NB_MAX / 64;
        std::vector<int32_t> st;
        if (d.g.chunkCost) {
            uint64_t lo = 0, hi = 0;
            for (int32_t c = 0; c < nC; c++) { lo = std::max(lo, d.g.chunkCost[c]); hi += d.g.chunkCost[c]; }
            while (lo < hi) {   // the smallest bound that packs into `need` ranges
                const uint64_t mid = lo + (hi - lo) / 2;
                if (pack_chunks(d.g.chunkCost, nC, CPB, mid, nullptr) <= need) hi = mid; else lo = mid + 1;
            }
            pack_chunks(d.g.chunkCost, nC, CPB, lo, &st);
            if ((int32_t)st.size() > need) st.clear();
        }
        if (st.empty())   // equal ranges
            for (int32_t r = 0; r < need; r++) st.push_back((int32_t)(r * nb / 64));
        for (int32_t &c : st) c *= 64;
        st.push_back(d.N);
        it = g_layouts.emplace(key, std::move(st)).first;
    }
    l.starts = it->second;
    l.nb = 0;
    for (size_t r = 0; r + 1 < l.starts.size(); r++) l.nb = std::max(l.nb, l.starts[r + 1] - l.starts[r]);
    l.S = (int32_t)l.starts.size() - 1 + l.G;
    if (l.S > S_MAX) return false;
    l.nSpare = std::min(D, maxB - l.S);
    l.grid = l.S + l.nSpare;
    return l.grid <= GBAR_MAX_BLOCKS;
}

static fb_status supported_locked(const DevBrain &d, bool serialDeliver, std::string *why) {
    if (serialDeliver) { *why = "presynaptic inhibition or tRef <= 0 needs the serial deliver path"; return FB_ERR_UNSUPPORTED; }
    if (d.k.bgEvents != 0.f) { *why = "background events draw from the shared stream every step"; return FB_ERR_UNSUPPORTED; }
    if (d.poolOf) { *why = "the pool loop sets the drive after every step"; return FB_ERR_UNSUPPORTED; }
    if (d.nDriven > DRV_CAP) { *why = "more than 8192 driven neurons"; return FB_ERR_UNSUPPORTED; }
    const int32_t D = d.nslots - 1;
    if (D < 1 || D > D_MAX) { *why = "delay outside 1..32 steps"; return FB_ERR_UNSUPPORTED; }
    if (d.k.depU != 0.f && refractory_steps(d.k) < D) {
        *why = "synaptic depression with a refractory period shorter than the delay";
        return FB_ERR_UNSUPPORTED;
    }
    if (KEY_THRESHOLD + (int64_t)d.N >= (1LL << 20)) { *why = "more than 1,040,383 neurons (arrival key field)"; return FB_ERR_UNSUPPORTED; }
    if (d.g.maxRow >= (1LL << 26)) { *why = "a row longer than 2^26 edges (row position field)"; return FB_ERR_UNSUPPORTED; }
    fb_status st = persist_init();
    if (st) { *why = "the device's occupancy query failed"; return st; }
    Layout l;
    if (!layout_for(d, D, l)) { *why = "more than 1024 neurons per workgroup at this grid size"; return FB_ERR_UNSUPPORTED; }
    const int32_t shareBlocks = std::min(l.grid - l.G, SHARE_BLOCKS);
    if (shareBlocks < 1 || (d.nDriven + shareBlocks - 1) / shareBlocks > SHQ) {
        *why = "too many driven neurons for this grid";
        return FB_ERR_UNSUPPORTED;
    }
    return FB_OK;
}

// (the layouts are cached in a map the launches also use)
fb_status persistent_supported(const DevBrain &d, bool serialDeliver, std::string *why) {
    std::lock_guard<std::mutex> lock(grid_launch_mutex());
    return supported_locked(d, serialDeliver, why);
}

struct PersistBuffers {
    uint64_t *spk = nullptr;
    uint32_t *dmask = nullptr, *tmask = nullptr;
    int32_t *drvPos = nullptr;
    uint32_t *arrMask = nullptr;
    uint64_t *bkKey = nullptr;
    float *bkAdd = nullptr;
    int32_t *bkCnt = nullptr;
    int32_t *ctl = nullptr;
    uint64_t *prof = nullptr;
    uint32_t *eown = nullptr;
    float *dres = nullptr;
    int32_t *bar = nullptr;
    uint64_t *xc = nullptr;
    int32_t *lrCnt = nullptr, *lrI = nullptr;
    uint32_t *lrK = nullptr;
    float *lrR = nullptr;
    uint32_t overflowSeen = 0;       // OR of the overflow flags of every launch (2 events, 4 spikes)
    float *dstage = nullptr;
    // The stage between launches (DSTAGE_F): its state fields hold the driven
    // neurons' state, its parameter fields the driven list's parameters; the
    // brain's arrays lag the stage until k_unstage runs; stageN neurons.
    bool stageState = false, stageParams = false, arraysBehind = false;
    int32_t stageN = 0;
    int32_t *hostCtl = nullptr;      // pinned: the last block's results (PersistArgs::hostCtl)
    uint64_t *hostProf = nullptr;
    int32_t tag = 0;                 // launches so far
    bool needClear = true;           // the counters a launch needs zero are not known to be
    int32_t N = 0, S = 0;
    int64_t E = 0;
    int device = -1;
    int32_t eownG = -1;
    std::vector<int32_t> starts;     // the regular ranges bstart / chunkBlk hold
    int32_t *bstart = nullptr;       // [S_MAX + 1]
    int32_t *chunkBlk = nullptr;     // [N/64] regular block of each 64-neuron chunk
    uint64_t eownDriven = ~0ull;
    // start and end events of the launches, two pairs used in turn: a
    // launch's time is taken after the next launch is queued, so that no
    // host work on the last launch's events comes before the next dispatch
    hipEvent_t ev[2][2] = {};
    int evSlot = 0;                  // the pair the next launch records
    bool timerPending[2] = {};       // the pair's time is not taken yet
};

// Adds the kernel time of event pair k (start .. end) to *kernelMs once its
// end event is done (waiting for it).
static hipError_t take_kernel_time(PersistBuffers &pb, int k, double *kernelMs) {
    if (!pb.timerPending[k]) return hipSuccess;
    pb.timerPending[k] = false;
    hipError_t e = hipEventSynchronize(pb.ev[k][1]);
    if (e != hipSuccess) return e;
    float ms = 0.f;
    e = hipEventElapsedTime(&ms, pb.ev[k][0], pb.ev[k][1]);
    if (e == hipSuccess && kernelMs) *kernelMs += ms;
    return e;
}

static std::map<const int32_t *, PersistBuffers> g_pbuf;   // keyed by the brain's driven buffer

static void free_buffers(PersistBuffers &pb) {
    void *ps[] = {pb.spk, pb.dmask, pb.tmask, pb.drvPos, pb.arrMask, pb.bkKey, pb.bkAdd, pb.bkCnt,
                  pb.ctl, pb.prof, pb.eown, pb.dres, pb.bar, pb.xc, pb.bstart, pb.chunkBlk,
                  pb.lrCnt, pb.lrI, pb.lrK, pb.lrR, pb.dstage};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    if (pb.hostCtl) (void)hipHostFree(pb.hostCtl);
    if (pb.hostProf) (void)hipHostFree(pb.hostProf);
    pb.hostCtl = nullptr;
    pb.hostProf = nullptr;
    pb.needClear = true;
    pb.spk = nullptr; pb.dmask = pb.tmask = nullptr; pb.drvPos = nullptr; pb.arrMask = nullptr;
    pb.bkKey = nullptr; pb.bkAdd = nullptr; pb.bkCnt = nullptr; pb.ctl = nullptr; pb.prof = nullptr;
    pb.eown = nullptr;
    pb.dres = nullptr;
    pb.bar = nullptr;
    pb.xc = nullptr;
    pb.lrCnt = pb.lrI = nullptr;
    pb.lrK = nullptr;
    pb.lrR = nullptr;
    pb.dstage = nullptr;
    pb.bstart = nullptr;
    pb.chunkBlk = nullptr;
    pb.starts.clear();
    pb.eownG = -1;
    pb.eownDriven = ~0ull;
}

fb_status launch_step(const DevBrain &d, int32_t head, int32_t slot, bool serialDeliver, bool rowsUnique,
                      hipStream_t st);

// Buffers of destroyed brains, kept for the next brain of the same shape on
// the same device (a brain per seed or per run would otherwise allocate and
// free ~100 MB each time).  The brain's stream is idle when it is released.
static std::vector<PersistBuffers> g_pool;
constexpr size_t POOL_MAX = 2;

void persistent_take_time(const int32_t *key, double *kernelMs) {
    std::lock_guard<std::mutex> lock(grid_launch_mutex());
    auto it = g_pbuf.find(key);
    if (it != g_pbuf.end())
        for (int k = 0; k < 2; k++) (void)take_kernel_time(it->second, k, kernelMs);
}

void persistent_release(const int32_t *key) {
    std::lock_guard<std::mutex> lock(grid_launch_mutex());
    auto it = g_pbuf.find(key);
    if (it == g_pbuf.end()) return;
    PersistBuffers pb = std::move(it->second);
    g_pbuf.erase(it);
    pb.timerPending[0] = pb.timerPending[1] = false;
    pb.stageState = pb.stageParams = pb.arraysBehind = false;   // the brain's state goes with it
    if (pb.spk && g_pool.size() < POOL_MAX) {
        pb.eownDriven = ~0ull;   // the next brain's driven set is its own
        pb.overflowSeen = 0;
        g_pool.push_back(std::move(pb));
        return;
    }
    free_buffers(pb);
    for (auto &pair : pb.ev)
        for (hipEvent_t &e : pair)
            if (e) (void)hipEventDestroy(e);
}

// The stage's driven neurons -> the brain's arrays (state and spike counts);
// the stage keeps the state, and its increments are cleared once added.
__global__ void k_unstage(DevBrain d, float *st, int32_t n) {
    const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int32_t i = __float_as_int(st[11 * DRV_CAP + q]);
    d.v[i] = st[0 * DRV_CAP + q]; d.gE[i] = st[1 * DRV_CAP + q]; d.gI[i] = st[2 * DRV_CAP + q];
    d.refr[i] = st[3 * DRV_CAP + q]; d.trace[i] = st[4 * DRV_CAP + q]; d.adapt[i] = st[5 * DRV_CAP + q];
    d.res[i] = st[6 * DRV_CAP + q];
    const uint32_t inc = __float_as_uint(st[12 * DRV_CAP + q]);
    if (inc) {
        d.spikeCount[i] += inc;
        st[12 * DRV_CAP + q] = 0.f;
    }
}

static hipError_t unstage_locked(const DevBrain &d, PersistBuffers &pb, hipStream_t st) {
    if (!pb.arraysBehind) return hipSuccess;
    pb.arraysBehind = false;
    if (pb.stageN > 0) hipLaunchKernelGGL(k_unstage, dim3((pb.stageN + 255) / 256), dim3(256), 0, st, d, pb.dstage, pb.stageN);
    return hipGetLastError();
}

fb_status persistent_sync(const DevBrain &d, hipStream_t st, bool state, bool params) {
    std::lock_guard<std::mutex> lock(grid_launch_mutex());
    auto it = g_pbuf.find(d.driven);
    if (it == g_pbuf.end()) return FB_OK;
    PersistBuffers &pb = it->second;
    FB_HIP(unstage_locked(d, pb, st));
    if (state) pb.stageState = false;
    if (params) pb.stageParams = false;
    return FB_OK;
}

__global__ void k_drvpos(int32_t *drvPos, int32_t N) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) drvPos[i] = -1;
}
__global__ void k_drvpos_set(int32_t *drvPos, const int32_t *driven, int32_t nD) {
    const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nD) drvPos[driven[q]] = q;
}

// the launch's control words, barrier counters, clocks, exchange words and
// bucket counts cleared in one kernel (five fills would be five queued
// operations)
__global__ void k_persist_clear(int32_t *ctl, int32_t *bar, uint64_t *prof, uint64_t *xc, int32_t *lrCnt, int32_t *bkCnt,
                                int32_t nCnt) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t n = gridDim.x * blockDim.x;
    if (i < 4) ctl[i] = 0;
    if (i < PERSIST_PROF) prof[i] = 0;
    if (i < DRV_BLOCKS * XC_RING) xc[i] = 0;
    for (int32_t k = i; k < LR_WORDS; k += n) lrCnt[k] = 0;
    for (int32_t k = i; k < GBAR_WORDS; k += n) bar[k] = 0;
    for (int32_t k = i; k < nCnt; k += n) bkCnt[k] = 0;
}

// destination block << 12 | local index, per edge: a driven target at list
// position p belongs to driven block p / Q (index p mod Q), others to block
// G + r for the regular range r that holds them
__global__ void k_edge_owner(uint32_t *eown, const uint2 *edges, int64_t E, const int32_t *drvPos,
                             const int32_t *chunkBlk, const int32_t *bstart, int32_t G, int32_t Q) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < E; j += (int64_t)gridDim.x * blockDim.x) {
        const int32_t q = (int32_t)edges[j].x;
        const int32_t dp = drvPos[q];
        const int32_t blk = chunkBlk[q >> 6];
        const int32_t db = dp >= 0 ? dp / Q : 0;
        eown[j] = dp >= 0 ? ((uint32_t)db << 12) | (uint32_t)(dp - db * Q)
                          : ((uint32_t)(G + blk) << 12) | (uint32_t)(q - bstart[blk]);
    }
}

std::mutex &grid_launch_mutex() {
    static std::mutex mu;
    return mu;
}

fb_status launch_persistent(const DevBrain &d, int32_t head, int32_t nsteps, hipStream_t st, int64_t *fallbackSteps,
                            int64_t *launches, double *kernelMs, double *phaseUs, uint64_t drivenVersion) {
    std::lock_guard<std::mutex> lock(grid_launch_mutex());
    std::string why;
    fb_status rc = supported_locked(d, false, &why);
    if (rc) { set_error("persistent engine: %s", why.c_str()); return rc; }
    const int32_t D = d.nslots - 1;
    Layout lay;
    layout_for(d, D, lay);
    const int32_t nwords = (d.N + 63) / 64;
    int dev = 0;
    FB_HIP(hipGetDevice(&dev));
    if (g_pbuf.find(d.driven) == g_pbuf.end()) {
        for (size_t k = 0; k < g_pool.size(); k++) {
            PersistBuffers &q = g_pool[k];
            if (q.N == d.N && q.S == lay.S && q.E == d.g.E && q.device == dev) {
                // a released brain's buffers: the bitmaps and arrival masks start clear
                FB_HIP(hipMemsetAsync(q.spk, 0, sizeof(uint64_t) * 2 * D_MAX * (size_t)nwords, st));
                FB_HIP(hipMemsetAsync(q.arrMask, 0, sizeof(uint32_t) * (size_t)d.N, st));
                g_pbuf[d.driven] = std::move(q);
                g_pool.erase(g_pool.begin() + (std::ptrdiff_t)k);
                break;
            }
        }
    }
    PersistBuffers &pb = g_pbuf[d.driven];
    if (pb.N != d.N || pb.S != lay.S || pb.E != d.g.E || pb.device != dev) {
        if (pb.device == dev) FB_HIP(unstage_locked(d, pb, st));   // the stage goes with the buffers
        pb.stageState = pb.stageParams = pb.arraysBehind = false;
        free_buffers(pb);
        pb.N = d.N;
        pb.S = lay.S;
        pb.E = d.g.E;
        pb.device = dev;
        FB_HIP(hipMalloc(&pb.spk, sizeof(uint64_t) * 2 * D_MAX * (size_t)nwords));
        FB_HIP(hipMemsetAsync(pb.spk, 0, sizeof(uint64_t) * 2 * D_MAX * (size_t)nwords, st));
        FB_HIP(hipMalloc(&pb.dmask, sizeof(uint32_t) * 2 * DRV_CAP));
        FB_HIP(hipMalloc(&pb.tmask, sizeof(uint32_t) * 2 * DRV_CAP));
        FB_HIP(hipMalloc(&pb.drvPos, sizeof(int32_t) * (size_t)d.N));
        FB_HIP(hipMalloc(&pb.arrMask, sizeof(uint32_t) * (size_t)d.N));
        FB_HIP(hipMemsetAsync(pb.arrMask, 0, sizeof(uint32_t) * (size_t)d.N, st));
        FB_HIP(hipMalloc(&pb.bkKey, sizeof(uint64_t) * 2 * (size_t)lay.S * EV_CAP));
        FB_HIP(hipMalloc(&pb.bkAdd, sizeof(float) * 2 * (size_t)lay.S * EV_CAP));
        FB_HIP(hipMalloc(&pb.bkCnt, sizeof(int32_t) * 2 * (size_t)lay.S * BK_STRIDE));
        FB_HIP(hipMalloc(&pb.ctl, sizeof(int32_t) * 4));
        FB_HIP(hipHostMalloc((void **)&pb.hostCtl, sizeof(int32_t) * 4, hipHostMallocDefault));
        FB_HIP(hipHostMalloc((void **)&pb.hostProf, sizeof(uint64_t) * PERSIST_PROF, hipHostMallocDefault));
        pb.hostCtl[2] = 0;
        pb.needClear = true;
        FB_HIP(hipMalloc(&pb.prof, sizeof(uint64_t) * PERSIST_PROF));
        FB_HIP(hipMalloc(&pb.eown, sizeof(uint32_t) * (size_t)std::max<int64_t>(d.g.E, 1)));
        FB_HIP(hipMalloc(&pb.dres, sizeof(float) * 2 * DRV_CAP));
        FB_HIP(hipMalloc(&pb.bar, sizeof(int32_t) * GBAR_WORDS));
        FB_HIP(hipMalloc(&pb.xc, sizeof(uint64_t) * DRV_BLOCKS * XC_RING));
        FB_HIP(hipMalloc(&pb.bstart, sizeof(int32_t) * (S_MAX + 1)));
        FB_HIP(hipMalloc(&pb.lrCnt, sizeof(int32_t) * LR_WORDS));
        FB_HIP(hipMalloc(&pb.dstage, sizeof(float) * DSTAGE_F * DRV_CAP));
        FB_HIP(hipMalloc(&pb.lrI, sizeof(int32_t) * 2 * LR_CAP));
        FB_HIP(hipMalloc(&pb.lrK, sizeof(uint32_t) * 2 * LR_CAP));
        FB_HIP(hipMalloc(&pb.lrR, sizeof(float) * 2 * LR_CAP));
        FB_HIP(hipMalloc(&pb.chunkBlk, sizeof(int32_t) * (size_t)((d.N + 63) / 64)));
        for (auto &pair : pb.ev)
            for (hipEvent_t &e : pair)
                if (!e) FB_HIP(hipEventCreate(&e));
    }
    // driven positions and edge owners (again when the driven set or layout changed)
    if (pb.starts != lay.starts) {
        // the regular ranges: block starts, and each chunk's block for the edge owners
        const int32_t R = (int32_t)lay.starts.size() - 1, nC = (d.N + 63) / 64;
        std::vector<int32_t> cb((size_t)nC);
        for (int32_t r = 0; r < R; r++)
            for (int32_t c = lay.starts[r] / 64; c < (lay.starts[r + 1] + 63) / 64; c++) cb[(size_t)c] = r;
        FB_HIP(hipMemcpyAsync(pb.bstart, lay.starts.data(), sizeof(int32_t) * (size_t)(R + 1), hipMemcpyHostToDevice, st));
        FB_HIP(hipMemcpyAsync(pb.chunkBlk, cb.data(), sizeof(int32_t) * (size_t)nC, hipMemcpyHostToDevice, st));
        FB_HIP(hipStreamSynchronize(st));   // the host vector goes out of scope
        pb.starts = lay.starts;
        pb.eownG = -1;
    }
    if (pb.eownDriven != drivenVersion || pb.eownG != lay.G) {
        // a new driven list or split: the stage's positions change
        FB_HIP(unstage_locked(d, pb, st));
        pb.stageState = pb.stageParams = false;
        hipLaunchKernelGGL(k_drvpos, dim3((d.N + 255) / 256), dim3(256), 0, st, pb.drvPos, d.N);
        if (d.nDriven)
            hipLaunchKernelGGL(k_drvpos_set, dim3((d.nDriven + 255) / 256), dim3(256), 0, st, pb.drvPos, d.driven,
                               d.nDriven);
        if (d.g.E)
            hipLaunchKernelGGL(k_edge_owner, dim3(1024), dim3(256), 0, st, pb.eown, d.g.edges, d.g.E, pb.drvPos,
                               pb.chunkBlk, pb.bstart, lay.G, std::max(lay.Q, 1));
        FB_HIP(hipGetLastError());
        pb.eownDriven = drivenVersion;
        pb.eownG = lay.G;
    }

    PersistArgs a{};
    a.d = d;
    a.D = D;
    a.nb = lay.nb;
    a.bstart = pb.bstart;
    a.G = lay.G;
    a.Q = lay.Q;
    a.S = lay.S;
    a.nSpare = lay.nSpare;
    a.nwords = nwords;
    a.spk = pb.spk;
    a.dmask = pb.dmask;
    a.tmask = pb.tmask;
    a.drvPos = pb.drvPos;
    a.arrMask = pb.arrMask;
    a.bkKey = pb.bkKey;
    a.bkAdd = pb.bkAdd;
    a.bkCnt = pb.bkCnt;
    a.ctl = pb.ctl;
    a.prof = pb.prof;
    a.eown = pb.eown;
    a.dres = pb.dres;
    a.bar = pb.bar;
    a.xc = pb.xc;
    a.lrCnt = pb.lrCnt;
    a.pflag = pb.lrCnt + 2;
    a.epc = pb.lrCnt + 64;
    a.lrI = pb.lrI;
    a.lrK = pb.lrK;
    a.lrR = pb.lrR;
    static const bool profileSteps = getenv("FB_PROFILE_DRIVEN_STEPS") != nullptr;

    a.hostCtl = pb.hostCtl;
    a.dstage = pb.dstage;
    a.hostProf = pb.hostProf;

    int32_t done = 0, cur = head;
    while (done < nsteps) {
        a.head0 = cur;
        a.nsteps = nsteps - done;
        a.tag = ++pb.tag;
        a.flags = (profileSteps ? 1 : 0) | (pb.stageState ? STAGE_STATE : 0) | (pb.stageParams ? STAGE_PARAMS : 0);
        if (pb.needClear) {
            hipLaunchKernelGGL(k_persist_clear, dim3(16), dim3(256), 0, st, pb.ctl, pb.bar, pb.prof, pb.xc, pb.lrCnt,
                               pb.bkCnt, 2 * lay.S * BK_STRIDE);
            FB_HIP(hipGetLastError());
            pb.needClear = false;
        }
        void *args[] = {&a};
        const int ek = pb.evSlot;
        FB_HIP(hipEventRecord(pb.ev[ek][0], st));
        // A plain launch: the grid is one workgroup per CU at the kernel's
        // occupancy and holds grid_launch_mutex (see there), which is what a
        // cooperative launch would guarantee; the cooperative launch path
        // cost ~13 us more per launch.
        const hipError_t le = hipLaunchKernel((const void *)k_persist, dim3(lay.grid), dim3(PT), args, 0, st);
        if (le != hipSuccess) {
            // Before any step ran this is reported as unsupported, so that the
            // automatic engine choice can step with the per-step engines.
            (void)hipGetLastError();
            set_error("persistent engine: launch refused: %s", hipGetErrorString(le));
            return done == 0 ? FB_ERR_UNSUPPORTED : FB_ERR_DEVICE;
        }
        // the driven neurons' state is in the stage from now on
        pb.stageState = pb.stageParams = pb.arraysBehind = true;
        pb.stageN = d.nDriven;
        FB_HIP(hipEventRecord(pb.ev[ek][1], st));
        pb.timerPending[ek] = true;
        pb.evSlot = ek ^ 1;
        // the previous launch's kernel time, now that this one is queued
        FB_HIP(take_kernel_time(pb, ek ^ 1, kernelMs));
        if (launches) (*launches)++;
        // Wait by polling: the last block out writes the tag to pinned memory
        // as its final act; a blocking stream synchronisation takes ~16 us
        // longer to return.  A kernel that ends without the tag (a block left
        // early) is caught by the stream query.  After the tag nothing of the
        // launch is left to wait for: the end event's time is taken at the
        // next launch or when the statistic is read (persistent_kernel_ms).
        for (uint32_t spin = 1;; spin++) {
            if (__atomic_load_n(&pb.hostCtl[2], __ATOMIC_ACQUIRE) == a.tag) break;
            if ((spin & 1023u) == 0) {
                const hipError_t q = hipStreamQuery(st);
                if (q == hipSuccess) break;
                if (q != hipErrorNotReady) FB_HIP(q);
            }
        }
        int32_t ctl[4];
        uint64_t prof[PERSIST_PROF];
        if (__atomic_load_n(&pb.hostCtl[2], __ATOMIC_ACQUIRE) == a.tag) {
            // the last block out wrote the results and cleared the counters
            ctl[0] = pb.hostCtl[0];
            ctl[1] = pb.hostCtl[1];
            memcpy(prof, pb.hostProf, sizeof prof);
        } else {
            // a block left early: read the device words, and clear before the next launch
            FB_HIP(hipMemcpy(ctl, pb.ctl, sizeof ctl, hipMemcpyDeviceToHost));
            FB_HIP(hipMemcpy(prof, pb.prof, sizeof prof, hipMemcpyDeviceToHost));
            pb.needClear = true;
        }
        if (phaseUs) {
            for (int k = 0; k < PROF_SPAN; k++)
                phaseUs[k] += (double)prof[k] / (k >= PROF_SHARE && k < PROF_SHARE + 3 ? std::min(lay.grid - lay.G, SHARE_BLOCKS) : 1) / g_ps.tickHz * 1e6;
            // the span: spread of the block starts, first start to last block
            // end, last block end to the final write (a slot stays zero when a
            // block left early)
            const uint64_t t0 = ~prof[PROF_SPAN];
            if (prof[PROF_SPAN] && prof[PROF_SPAN + 3]) {
                phaseUs[PROF_SPAN] += (double)(prof[PROF_SPAN + 1] - t0) / g_ps.tickHz * 1e6;
                phaseUs[PROF_SPAN + 1] += (double)(prof[PROF_SPAN + 2] - t0) / g_ps.tickHz * 1e6;
                phaseUs[PROF_SPAN + 2] += (double)(prof[PROF_SPAN + 3] - prof[PROF_SPAN + 2]) / g_ps.tickHz * 1e6;
                for (int k = 4; k < 7; k++)
                    if (prof[PROF_SPAN + k]) phaseUs[PROF_SPAN + k] += (double)(prof[PROF_SPAN + k] - t0) / g_ps.tickHz * 1e6;
            }
        }
        if (ctl[1] & GBAR_TIMED_OUT) {
            set_error("persistent engine: a grid barrier timed out; the brain state is undefined");
            return FB_ERR_DEVICE;
        }
        done += ctl[0];
        pb.overflowSeen |= (uint32_t)ctl[1];
        // a stop can leave bucket counts behind (an overflow in the prologue's
        // push skips every gather): clear before the next launch
        if (ctl[1]) pb.needClear = true;
        cur = (cur + ctl[0]) % d.nslots;
        if (done < nsteps) {
            // a bounded buffer overflowed: the next epoch runs on the per-step
            // engine, on the brain's arrays
            FB_HIP(unstage_locked(d, pb, st));
            pb.stageState = false;
            if (d.k.depU != 0.f)
                FB_HIP(hipMemsetAsync(pb.arrMask, 0, sizeof(uint32_t) * (size_t)d.N, st));
            const int32_t m = std::min(D, nsteps - done);
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

uint32_t persistent_overflow_flags(const DevBrain &d) {
    std::lock_guard<std::mutex> lock(grid_launch_mutex());
    auto it = g_pbuf.find(d.driven);
    return it == g_pbuf.end() ? 0u : it->second.overflowSeen;
}

}  // namespace fb
