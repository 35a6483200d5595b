// fb_batch.hip -- B ensemble members of one connectome in one engine.
//
// The reference runs one brain per worker over a shared read-only graph
// (src/brainsetup.js:8-27) and its Bend port runs one member per lane
// (kernel.bend:1-17, 293-378).  Here member b of up to 64 is lane b of a
// wavefront: per-member state is laid out [N][B], so one wave advances one
// neuron in every member with coalesced loads, and a step's spikes are one
// 64-bit ballot per neuron.  Each member is exactly a single brain with
//   seed            its own ((seed * 2654435761) >>> 0 || 1 xorshift32 stream)
//   weights         W[j] = f32(count[j] * gain_b[type(pre)])  (writeGraph's
//                   outScale rule, src/lifwasm.js:23,31, with a per-type gain)
//   thr             its own per-neuron threshold offsets (silencing is
//                   thr = 1e6, scripts/lif_ensemble.mjs:35)
//   run length      it stops (state frozen, no draws) after its own step count
// and all members share the driven list and drive rates.
//
// Per step:
//   kb_deliver   every presynaptic neuron with a nonzero arrival mask walks
//                its row once (lanes = edges); for each set bit b the edge
//                events go into member b's per-(sign, target) accumulators with
//                the ordered-sum scheme of the single brain (bitmask x CSR).
//   kb_driven    member b's background and forced draws: per-wave counts of
//                non-refractory driven neurons, a prefix across waves per lane,
//                and a leapfrog jump of member b's stream to its offset.
//   kb_membrane  wave = neuron, lane = member: arrivals, membrane, threshold,
//                ballot; neurons with any spike are appended to the slot's
//                active list.
#include <algorithm>
#include <map>
#include <tuple>
#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

#include "fb_batch.h"

namespace fb {

uint32_t seed_hash(double seed);
void derive_consts(const fb_params &p, int32_t N, Consts &c);
int32_t nslots_of(const fb_params &p);

__device__ __forceinline__ uint64_t lanes_active(const BatchDev &d, int32_t t) {
    // members still running at launch step t
    const int lane = threadIdx.x & 63;
    const bool on = lane < d.B && t < d.steps[lane];
    return __ballot(on);
}

// ------------------------------------------------------------------ deliver

__global__ void __launch_bounds__(256) kb_deliver(BatchDev d, int32_t head, int32_t t) {
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    const uint64_t on = lanes_active(d, t);
    const int32_t A = d.nActive[head];
    const int32_t *list = d.active + (int64_t)head * d.N;
    const int32_t N = d.N, B = d.B;
    unsigned long long waveEdges = 0, waveSyn = 0;
    for (int32_t k = wave; k < A; k += nwaves) {
        const int32_t pre = list[k];
        uint64_t bits = d.mask[(int64_t)head * N + pre] & on;
        const uint64_t forced = d.fmask[(int64_t)head * N + pre];
        const uint32_t ra = d.g.indptr[pre], re = d.g.indptr[pre + 1];
        const float sgn = d.g.sign[pre];
        const int32_t typ = d.typeOf[pre];
        const int32_t fr = forced ? d.frank[(int64_t)head * N + pre] : 0;
        // lane b holds member b's gain for this row's type and its res
        const float gLane = lane < B ? d.gain[(int64_t)lane * d.nTypes + typ] : 0.f;
        const float rLane = (d.liveRes && lane < B) ? d.res[(int64_t)pre * B + lane] : 1.f;
        while (bits) {
            const int b = __ffsll((unsigned long long)bits) - 1;
            bits &= bits - 1;
            const int64_t ib = (int64_t)pre * B + b;
            const float resb = __shfl(rLane, b, 64);
            const float s = d.liveRes ? sgn * resb : sgn;
            if (s == 0.f) continue;
            // arrival rank in member b's lif.c list: forced spikes (driven
            // order) before threshold spikes (ascending index)
            const uint32_t rank = ((forced >> b) & 1ull) ? (uint32_t)fr : (uint32_t)N + (uint32_t)pre;
            if (lane == 0) {
                if (d.liveRes) d.res[ib] = resb - d.k.depU * resb;
                waveEdges += (unsigned long long)(re - ra);
                waveSyn += (unsigned long long)d.g.rowSyn[pre];
            }
            const float g = __shfl(gLane, b, 64);
            const int32_t cls = s > 0 ? 0 : N;
            const float sg = s * g;                  // lane mode: the gain scales the sign
            for (uint32_t j = ra + lane; j < re; j += 64) {
                const uint32_t q = d.g.edges[j].x;
                const float addend = d.signGain ? __uint_as_float(d.g.edges[j].y) * sg : ((float)d.g.counts[j] * g) * s;
                const int64_t ci = ((int64_t)cls + q) * B + b;
                const uint64_t kk = ((uint64_t)rank << 32) | (uint64_t)(j - ra);
                const int32_t slot = atomicAdd(&d.cnt[ci], 1);
                if (slot == 0) {
                    d.add[ci] = addend;
                    d.key[ci] = kk;
                } else {
                    const int32_t p = atomicAdd(d.ncoll, 1);
                    if (p < d.collCap) {
                        d.coll[p].key = kk;
                        d.coll[p].addend = addend;
                        d.coll[p].next = atomicExch(&d.chain[ci], p);
                    } else {
                        atomicOr(d.err, FB_DEVERR_COLL_OVERFLOW);
                    }
                }
            }
        }
    }
    if (lane == 0 && waveEdges) {
        atomicAdd(&d.counters[0], waveEdges);
        atomicAdd(&d.counters[1], waveSyn);
    }
}

// ------------------------------------------------------------------ driven

__global__ void __launch_bounds__(1024) kb_driven(BatchDev d, int32_t slot, int32_t t) {
    constexpr int T = 1024, W = T / 64;
    __shared__ uint32_t leapS[LEAP_LDS_WORDS];
    __shared__ int32_t cnt[W][64];
    __shared__ int32_t s_nbg;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    leap_load<T>(leapS, d.leap);
    if (tid == 0) {
        float acc = d.bgAcc[0] + d.k.bgEvents;
        int32_t nbg = (int32_t)acc;
        acc = acc - (float)nbg;
        d.bgAcc[0] = acc;
        // the leapfrog tables reach 2047 draws per step, background and forced together
        if (nbg + d.nDriven > 2047) { atomicOr(d.err, 2); nbg = 0; }
        s_nbg = nbg;
    }
    __syncthreads();
    const bool on = lane < d.B && t < d.steps[lane];
    uint32_t x0 = on ? d.rng[lane] : 0u;
    const int32_t nbg = s_nbg;
    // background: draws 1..nbg of each running member's stream (lif.c:48);
    // wave w takes draws w+1, w+1+W, ... via x -> M^(p) x
    if (on && nbg > 0) {
        for (int32_t p = wave + 1; p <= nbg; p += W) {
            const uint32_t y = matvec(leapS + (32 + (p >> 5)) * LEAP_STRIDE, x0);   // M^(32 (p/32)) x
            const uint32_t z = (p & 31) ? matvec(leapS + ((p & 31) - 1) * LEAP_STRIDE, y) : y;
            atomicAdd(&d.bgCnt[(int64_t)(z % (uint32_t)d.N) * d.B + lane], 1);
        }
        // state after nbg draws
        const uint32_t y = matvec(leapS + (32 + (nbg >> 5)) * LEAP_STRIDE, x0);
        x0 = (nbg & 31) ? matvec(leapS + ((nbg & 31) - 1) * LEAP_STRIDE, y) : y;
    }
    // forced spikes: wave w owns driven positions [lo, hi) for every member.
    // The slice is walked CH neurons at a time with every load of a chunk in
    // flight together; the draws themselves then run from registers.
    constexpr int CH = 32;
    const int32_t nD = d.nDriven;
    const int32_t per = (nD + W - 1) / W;
    const int32_t lo = min(nD, wave * per), hi = min(nD, lo + per);
    const int32_t B = d.B;
    int32_t mine = 0;
    if (on)
        for (int32_t c0 = lo; c0 < hi; c0 += CH) {
            float r[CH];
#pragma unroll
            for (int u = 0; u < CH; u++)
                r[u] = c0 + u < hi ? d.refr[(int64_t)d.driven[c0 + u] * B + lane] : 1.f;
#pragma unroll
            for (int u = 0; u < CH; u++) mine += !(r[u] > 0) ? 1 : 0;
        }
    cnt[wave][lane] = mine;
    __syncthreads();
    int32_t off = 0, total = 0;
    for (int w = 0; w < W; w++) {
        const int32_t c = cnt[w][lane];
        off += w < wave ? c : 0;
        total += c;
    }
    // the state before this wave's first draw: M^off x0 (off < 2048)
    uint32_t x = x0;
    if (on && mine > 0) {
        const uint32_t y = matvec(leapS + (32 + (off >> 5)) * LEAP_STRIDE, x0);
        x = (off & 31) ? matvec(leapS + ((off & 31) - 1) * LEAP_STRIDE, y) : y;
    }
    // a forced spike's effects on the neuron (lif.c:51-52) are applied by
    // kb_membrane from the mask, so this kernel reads state and writes masks only
    for (int32_t c0 = lo; c0 < hi; c0 += CH) {
        float r[CH], pf[CH];
#pragma unroll
        for (int u = 0; u < CH; u++) {
            const bool in = on && c0 + u < hi;
            const int32_t i = c0 + u < hi ? d.driven[c0 + u] : 0;
            r[u] = in ? d.refr[(int64_t)i * B + lane] : 1.f;
            pf[u] = in ? (d.laneDrive ? d.laneDrive[(int64_t)i * B + lane] : d.drive[i]) * d.k.dtS : 0.f;
        }
#pragma unroll
        for (int u = 0; u < CH; u++) {
            if (c0 + u >= hi) break;
            const int32_t i = d.driven[c0 + u];
            bool f = false;
            if (!(r[u] > 0)) {
                x = xs32(x);
                const float uu = (float)(x >> 8) * (1.0f / 16777216.0f);
                f = uu < pf[u];
            }
            const uint64_t fb = __ballot(f);
            if (lane == 0) {
                d.fmask[(int64_t)slot * d.N + i] = fb;
                d.frank[(int64_t)slot * d.N + i] = c0 + u;
            }
        }
    }
    if (wave == 0 && on) {
        // the stream after all of this step's draws
        const uint32_t y = matvec(leapS + (32 + (total >> 5)) * LEAP_STRIDE, x0);
        d.rng[lane] = (total & 31) ? matvec(leapS + ((total & 31) - 1) * LEAP_STRIDE, y) : y;
    }
}

// The draws over DRV_WG workgroups when there are no background events:
// kb_driven_count writes each wave's non-refractory count per lane (its
// slice of the driven list), then kb_driven_multi starts every wave at its
// stream position (the counts of the waves before it) with the leapfrog
// jump, draws its slice as kb_driven does, and leaves the state after the
// step in rngNext for kb_membrane to install (every wave reads rng first).
__global__ void __launch_bounds__(1024) kb_driven_count(BatchDev d, int32_t t) {
    constexpr int W = 16, CH = 32;
    const int tid = threadIdx.x, lane = tid & 63, gw = blockIdx.x * W + (tid >> 6);
    const bool on = lane < d.B && t < d.steps[lane];
    const int32_t nD = d.nDriven, GW = gridDim.x * W;
    const int32_t per = (nD + GW - 1) / GW;
    const int32_t lo = min(nD, gw * per), hi = min(nD, lo + per);
    int32_t mine = 0;
    if (on)
        for (int32_t c0 = lo; c0 < hi; c0 += CH) {
            float r[CH];
#pragma unroll
            for (int u = 0; u < CH; u++) r[u] = c0 + u < hi ? d.refr[(int64_t)d.driven[c0 + u] * d.B + lane] : 1.f;
#pragma unroll
            for (int u = 0; u < CH; u++) mine += !(r[u] > 0) ? 1 : 0;
        }
    d.drvPart[gw * 64 + lane] = mine;
}

__global__ void __launch_bounds__(1024) kb_driven_multi(BatchDev d, int32_t slot, int32_t t) {
    constexpr int T = 1024, W = T / 64, CH = 32;
    __shared__ uint32_t leapS[LEAP_LDS_WORDS];
    const int tid = threadIdx.x, lane = tid & 63, gw = blockIdx.x * W + (tid >> 6);
    leap_load<T>(leapS, d.leap);
    __syncthreads();
    const bool on = lane < d.B && t < d.steps[lane];
    const uint32_t x0 = on ? d.rng[lane] : 0u;
    const int32_t nD = d.nDriven, GW = gridDim.x * W, B = d.B;
    const int32_t per = (nD + GW - 1) / GW;
    const int32_t lo = min(nD, gw * per), hi = min(nD, lo + per);
    int32_t off = 0, total = 0;
    for (int32_t w = 0; w < GW; w++) {
        const int32_t c = d.drvPart[w * 64 + lane];
        off += w < gw ? c : 0;
        total += c;
    }
    uint32_t x = x0;
    if (on && off > 0) {
        const uint32_t y = matvec(leapS + (32 + (off >> 5)) * LEAP_STRIDE, x0);
        x = (off & 31) ? matvec(leapS + ((off & 31) - 1) * LEAP_STRIDE, y) : y;
    }
    for (int32_t c0 = lo; c0 < hi; c0 += CH) {
        float r[CH], pf[CH];
#pragma unroll
        for (int u = 0; u < CH; u++) {
            const bool in = on && c0 + u < hi;
            const int32_t i = c0 + u < hi ? d.driven[c0 + u] : 0;
            r[u] = in ? d.refr[(int64_t)i * B + lane] : 1.f;
            pf[u] = in ? (d.laneDrive ? d.laneDrive[(int64_t)i * B + lane] : d.drive[i]) * d.k.dtS : 0.f;
        }
#pragma unroll
        for (int u = 0; u < CH; u++) {
            if (c0 + u >= hi) break;
            const int32_t i = d.driven[c0 + u];
            bool f = false;
            if (!(r[u] > 0)) {
                x = xs32(x);
                const float uu = (float)(x >> 8) * (1.0f / 16777216.0f);
                f = uu < pf[u];
            }
            const uint64_t fb = __ballot(f);
            if (lane == 0) {
                d.fmask[(int64_t)slot * d.N + i] = fb;
                d.frank[(int64_t)slot * d.N + i] = c0 + u;
            }
        }
    }
    if (gw == 0 && on) {
        const uint32_t y = matvec(leapS + (32 + (total >> 5)) * LEAP_STRIDE, x0);
        d.rngNext[lane] = (total & 31) ? matvec(leapS + ((total & 31) - 1) * LEAP_STRIDE, y) : y;
    }
}

// ------------------------------------------------------------------ membrane + threshold

__global__ void __launch_bounds__(256) kb_membrane(BatchDev d, int32_t slot, int32_t head, int32_t t) {
    const int lane = threadIdx.x & 63;
    // the multi-workgroup draws left the stream state after the step
    if (d.multiDrv && blockIdx.x == 0 && threadIdx.x < d.B && t < d.steps[threadIdx.x])
        d.rng[threadIdx.x] = d.rngNext[threadIdx.x];
    const int32_t i = (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    if (i >= d.N) return;   // whole wave
    const int32_t B = d.B, N = d.N;
    const bool on = lane < B && t < d.steps[lane];
    bool spk = false;
    const bool drv = d.drvPos[i] >= 0;
    const uint64_t fm = drv ? d.fmask[(int64_t)slot * N + i] : 0ull;
    if (on) {
        const int64_t ib = (int64_t)i * B + lane;
        NeuronState s;
        s.v = d.v[ib]; s.gE = d.gE[ib]; s.gI = d.gI[ib]; s.refr = d.refr[ib];
        s.trace = d.trace[ib];
        s.adapt = d.liveAdapt ? d.adapt[ib] : 0.f;
        s.res = d.liveRes ? d.res[ib] : 1.f;
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const int64_t ci = ((int64_t)c * N + i) * B + lane;
            const int32_t n = d.cnt[ci];
            if (n) {
                float g = c ? s.gI : s.gE;
                if (n == 1) g = g + d.add[ci];
                else {
                    g = ordered_sum_chain(g, n, d.key[ci], d.add[ci], d.chain[ci], d.coll);
                    d.chain[ci] = -1;
                }
                d.cnt[ci] = 0;
                if (c) s.gI = g; else s.gE = g;
            }
        }
        const int32_t nb = d.liveBg ? d.bgCnt[ib] : 0;
        if (nb) {
            for (int32_t q = 0; q < nb; q++) s.gE = s.gE + d.k.bgAmp;
            d.bgCnt[ib] = 0;
        }
        if ((fm >> lane) & 1ull) {
            // this step's forced spike (lif.c:51-52), after every delivery
            s.v = d.k.vReset;
            s.refr = d.k.tRef + d.k.dt;
            s.trace = 1.f;
            s.adapt = s.adapt + d.k.adaptInc;
            d.spikeCount[ib] += 1u;
        }
        membrane(d.k, s, d.bias[ib]);
        if (crosses(d.k, s, d.thr[ib])) {
            spike_effects(d.k, s, d.k.tRef);
            d.spikeCount[ib] += 1u;
            spk = true;
        }
        d.v[ib] = s.v; d.gE[ib] = s.gE; d.gI[ib] = s.gI; d.refr[ib] = s.refr;
        d.trace[ib] = s.trace;
        if (d.liveAdapt) d.adapt[ib] = s.adapt;
        if (d.liveRes) d.res[ib] = s.res;
    }
    // this step's arrivals are consumed: their slot becomes the next step's
    // output slot, and the chain pool is free (its entries were read above)
    if (i == 0 && lane == 0) {
        d.nActive[head] = 0;
        d.ncoll[0] = 0;
    }
    const uint64_t th = __ballot(spk);
    if (lane == 0) {
        if (!drv) d.fmask[(int64_t)slot * N + i] = 0ull;
        const uint64_t m = th | fm;
        d.mask[(int64_t)slot * N + i] = m;
        if (m) {
            const int32_t p = atomicAdd(&d.nActive[slot], 1);
            d.active[(int64_t)slot * N + p] = i;
            atomicAdd(&d.counters[2], (unsigned long long)__popcll(m));
        }
    }
}

// The per-lane pool loop (cordens.bend:227-273) after a step: each lane's
// spikes of the step tallied by pool (from the pool cells' masks), the
// counts recorded, and each lane's drive of the loop's targets for the next
// step set to base + gain * (its pools' total loopDelay[b] steps back), in
// binary32 (0 before the first step).
__global__ void __launch_bounds__(256) kb_pool_loop(BatchDev d, int32_t slot) {
    __shared__ uint32_t cnt[BMAX][8];
    __shared__ float s_drive[BMAX];
    const int tid = threadIdx.x;
    for (int q = tid; q < BMAX * 8; q += 256) cnt[q >> 3][q & 7] = 0u;
    __syncthreads();
    // every neuron's mask of the step is written (zero when silent), so the
    // pools' own cells are read rather than the step's whole list
    for (int32_t q = tid; q < d.loopNCells; q += 256) {
        const int32_t i = d.loopCells[q];
        const uint32_t p = d.poolOf[i];
        uint64_t m = d.mask[(int64_t)slot * d.N + i];
        while (m) {
            const int b = __ffsll((unsigned long long)m) - 1;
            m &= m - 1;
            atomicAdd(&cnt[b][p], 1u);
        }
    }
    __syncthreads();
    const int32_t k = d.loopState[0] + 1, r = d.loopState[1];
    if (tid < d.B) {
        const int b = tid;
        uint32_t total = 0;
        for (int32_t l = 0; l < d.loopPools; l++) {
            total += cnt[b][l];
            if (r < d.loopRecCap)
                d.loopRec[((int64_t)b * d.loopRecCap + r) * d.loopPools + l] = (uint16_t)min(cnt[b][l], 65535u);
        }
        uint32_t *h = d.loopHist + (int64_t)b * d.loopRing;
        h[k % d.loopRing] = total;
        const int32_t dl = d.loopDelay[b];
        const uint32_t back = k - dl >= 1 ? h[(k - dl) % d.loopRing] : 0u;
        const float gb = d.loopGain * (float)back;
        s_drive[b] = d.loopBase + gb;
    }
    __syncthreads();
    if (tid == 0) {
        d.loopState[0] = k;
        if (r < d.loopRecCap) d.loopState[1] = r + 1;
    }
    const int32_t B = d.B;
    for (int32_t q = tid; q < d.loopNTargets * B; q += 256) {
        const int32_t t = q / B, b = q - t * B;
        d.laneDrive[(int64_t)d.loopTargets[t] * B + b] = s_drive[b];
    }
}

// records the active list, spike masks and forced masks of one slot
__global__ void kb_record(BatchDev d, int32_t slot, int32_t *cntOut, int32_t *ids, uint64_t *masks, uint64_t *fmasks,
                          int32_t cap) {
    const int32_t n = d.nActive[slot];
    if (blockIdx.x == 0 && threadIdx.x == 0) *cntOut = n;
    for (int32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < min(n, cap); k += gridDim.x * blockDim.x) {
        const int32_t i = d.active[(int64_t)slot * d.N + k];
        ids[k] = i;
        masks[k] = d.mask[(int64_t)slot * d.N + i];
        fmasks[k] = d.fmask[(int64_t)slot * d.N + i];
    }
}

__global__ void kb_init(BatchDev d) {
    const int64_t n = (int64_t)d.N * d.B;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        d.v[q] = d.k.vRest; d.gE[q] = 0.f; d.gI[q] = 0.f; d.refr[q] = 0.f; d.trace[q] = 0.f;
        d.adapt[q] = 0.f; d.res[q] = 1.f; d.spikeCount[q] = 0u; d.bias[q] = 0.f; d.thr[q] = 0.f;
        d.bgCnt[q] = 0;
        d.cnt[q] = 0; d.cnt[q + n] = 0; d.chain[q] = -1; d.chain[q + n] = -1;
    }
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < (int64_t)d.nslots * d.N;
         q += (int64_t)gridDim.x * blockDim.x) {
        d.mask[q] = 0ull;
        d.fmask[q] = 0ull;
    }
    if (blockIdx.x == 0 && threadIdx.x < d.nslots) d.nActive[threadIdx.x] = 0;
}

__global__ void kb_set_member(float *arr, int32_t B, int32_t b, const int32_t *idx, int32_t n, float value) {
    const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) arr[(int64_t)idx[q] * B + b] = value;
}

}  // namespace fb

using namespace fb;

struct fb_batch {
    fb_graph *g = nullptr;
    fb_params p{};
    BatchDev d{};
    hipStream_t stream = nullptr, ownStream = nullptr;
    std::vector<void *> allocs;
    std::vector<int32_t> drivenList;
    std::vector<int32_t> steps;
    std::vector<int32_t> typeOf;
    int32_t head = 0, lastSlot = 0, t = 0;
    int32_t *idxScratch = nullptr;
    float *laneDriveBuf = nullptr;   // kept when lane drive is switched off, reused when it is set again
    int32_t *recCnt = nullptr, *recIds = nullptr;
    uint64_t *recMasks = nullptr, *recF = nullptr;
    int32_t engine = FB_ENGINE_AUTO;
    uint64_t drivenVersion = 0;
    int64_t laLaunches = 0, laFallback = 0, laSteps = 0;
    double laMs = 0.0;
    double laPhaseUs[10] = {0};      // ensemble lookahead clocks: driven then regular workgroup, 5 phases each
    // the lookahead engine's regular-block shape: chosen by the autotuner
    // (laShape, -1 until measured) or forced by the caller (laForced)
    int32_t laShape = -1, laForced = -1;
    double tuneMs[LA_NSHAPES] = {};
    int32_t tuneEpochs[LA_NSHAPES] = {};
    struct Trial {
        int32_t shape, epochs;
        bool timed;
    };
    std::vector<Trial> tunePlan;      // launches still to run before the choice
};

namespace fb {
bool batch_la_supported(const BatchDev &d, std::string *why);
bool batch_la_shape_fits(const BatchDev &d, int k);
int32_t batch_la_shape_neurons(int k);
void batch_la_release(const void *key);
fb_status batch_la_launch(const BatchDev &d, int k, int32_t t0, int32_t head0, int32_t epochs,
                          uint64_t drivenVersion, uint64_t *recM, uint64_t *recF, hipStream_t st, int32_t *done,
                          double *kernelMs, double *phaseUs);
}

// graph internals needed here (defined in fb_api.hip)
struct fb_graph;
extern "C" fb_status fb_graph_retain(fb_graph *g);
extern "C" fb_status fb_graph_release(fb_graph *g);
namespace fb {
const DevGraph *graph_dev(const fb_graph *g);
fb_status leap_tables(uint32_t **jump, uint32_t **leap);
}

template <class T>
static fb_status balloc(fb_batch *b, T **p, size_t count) {
    void *q = nullptr;
    FB_HIP(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
    b->allocs.push_back(q);
    *p = (T *)q;
    return FB_OK;
}

#define FB_REQUIRE_B(cond, code, ...)    \
    do {                                 \
        if (!(cond)) {                   \
            set_error(__VA_ARGS__);      \
            return code;                 \
        }                                \
    } while (0)

extern "C" fb_status fb_batch_create(fb_graph *g, const fb_params *p, int32_t B, const double *seeds,
                                     const int32_t *typeOf, int32_t nTypes, const float *gains, fb_batch **out) {
    FB_REQUIRE_B(g && p && seeds && typeOf && gains && out, FB_ERR_ARG, "fb_batch_create: null argument");
    FB_REQUIRE_B(B >= 1 && B <= BMAX, FB_ERR_ARG, "fb_batch_create: 1 <= B <= 64");
    FB_REQUIRE_B(nTypes >= 1, FB_ERR_ARG, "fb_batch_create: nTypes >= 1");
    const DevGraph *dg = graph_dev(g);
    FB_REQUIRE_B(dg->counts, FB_ERR_ARG, "fb_batch_create: the graph needs synapse counts");
    FB_REQUIRE_B(p->preInh == 0.0, FB_ERR_UNSUPPORTED, "fb_batch_create: presynaptic inhibition is not batched");
    FB_REQUIRE_B(p->tRef > 0.0, FB_ERR_UNSUPPORTED, "fb_batch_create: tRef must be > 0");
    const int32_t N = dg->N;
    for (int32_t i = 0; i < N; i++)
        FB_REQUIRE_B(typeOf[i] >= 0 && typeOf[i] < nTypes, FB_ERR_ARG, "fb_batch_create: type of neuron %d out of range", i);
    fb_batch *b = new fb_batch();
    b->g = g;
    fb_graph_retain(g);
    b->p = *p;
    BatchDev &d = b->d;
    d.N = N; d.B = B; d.nslots = nslots_of(*p); d.nDriven = 0; d.nTypes = nTypes;
    derive_consts(*p, N, d.k);
    d.liveAdapt = d.k.adaptInc != 0.f;
    d.liveRes = d.k.depU != 0.f;
    d.liveBg = d.k.bgEvents > 0.f;
    d.g = *dg;
    FB_HIP(hipStreamCreateWithFlags(&b->ownStream, hipStreamNonBlocking));
    b->stream = b->ownStream;
    const size_t NB = (size_t)N * B;
    fb_status st;
    float *f;
    if ((st = balloc(b, &f, NB * 9))) return st;
    d.v = f; d.gE = f + NB; d.gI = f + 2 * NB; d.refr = f + 3 * NB; d.trace = f + 4 * NB; d.adapt = f + 5 * NB;
    d.res = f + 6 * NB; d.bias = f + 7 * NB; d.thr = f + 8 * NB;
    float *drive; int32_t *driven, *drvPos, *typ, *stepsD, *active, *nActive, *cnt, *chain, *ncoll, *bgCnt, *err;
    float *gain, *add, *bgAcc; uint32_t *rng, *sc; uint64_t *mask, *fmask, *key; CollEntry *coll;
    unsigned long long *counters;
    if ((st = balloc(b, &sc, NB)) || (st = balloc(b, &drive, (size_t)N)) || (st = balloc(b, &driven, (size_t)N)) ||
        (st = balloc(b, &drvPos, (size_t)N)) || (st = balloc(b, &typ, (size_t)N)) ||
        (st = balloc(b, &gain, (size_t)B * nTypes)) || (st = balloc(b, &rng, (size_t)B)) ||
        (st = balloc(b, &bgAcc, 1)) || (st = balloc(b, &stepsD, (size_t)B)) ||
        (st = balloc(b, &mask, (size_t)d.nslots * N)) || (st = balloc(b, &fmask, (size_t)d.nslots * N)) ||
        (st = balloc(b, &d.frank, (size_t)d.nslots * N)) ||
        (st = balloc(b, &active, (size_t)d.nslots * N)) || (st = balloc(b, &nActive, (size_t)d.nslots)) ||
        (st = balloc(b, &cnt, 2 * NB)) || (st = balloc(b, &chain, 2 * NB)) || (st = balloc(b, &add, 2 * NB)) ||
        (st = balloc(b, &key, 2 * NB)) || (st = balloc(b, &bgCnt, NB)) || (st = balloc(b, &ncoll, 1)) ||
        (st = balloc(b, &err, 1)) || (st = balloc(b, &counters, 4)) || (st = balloc(b, &b->idxScratch, (size_t)N)) ||
        (st = balloc(b, &d.drvPart, (size_t)DRV_WG * 16 * 64)) || (st = balloc(b, &d.rngNext, (size_t)BMAX)))
        return st;
    d.collCap = std::max<int64_t>(dg->E * 4, 1);
    if ((st = balloc(b, &coll, (size_t)d.collCap))) return st;
    d.spikeCount = sc; d.drive = drive; d.driven = driven; d.drvPos = drvPos; d.typeOf = typ; d.gain = gain;
    d.rng = rng; d.bgAcc = bgAcc; d.steps = stepsD; d.mask = mask; d.fmask = fmask; d.active = active;
    d.nActive = nActive; d.cnt = cnt; d.chain = chain; d.add = add; d.key = key; d.coll = coll; d.ncoll = ncoll;
    d.bgCnt = bgCnt; d.err = err; d.counters = counters;
    uint32_t *jump = nullptr, *leap = nullptr;
    if ((st = leap_tables(&jump, &leap))) return st;
    d.leap = leap;
    std::vector<uint32_t> r((size_t)B);
    for (int32_t q = 0; q < B; q++) r[q] = seed_hash(seeds[q]);
    FB_HIP(hipMemcpy(rng, r.data(), sizeof(uint32_t) * B, hipMemcpyHostToDevice));
    FB_HIP(hipMemcpy(typ, typeOf, sizeof(int32_t) * N, hipMemcpyHostToDevice));
    FB_HIP(hipMemcpy(gain, gains, sizeof(float) * (size_t)B * nTypes, hipMemcpyHostToDevice));
    {
        float *gT = nullptr;
        if ((st = balloc(b, &gT, (size_t)B * nTypes))) return st;
        std::vector<float> t((size_t)B * nTypes);
        for (int32_t q = 0; q < B; q++)
            for (int32_t T = 0; T < nTypes; T++) t[(size_t)T * B + q] = gains[(size_t)q * nTypes + T];
        FB_HIP(hipMemcpy(gT, t.data(), sizeof(float) * t.size(), hipMemcpyHostToDevice));
        d.gainT = gT;
    }
    std::vector<int32_t> neg((size_t)N, -1);
    FB_HIP(hipMemcpy(drvPos, neg.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice));
    FB_HIP(hipMemset(drive, 0, sizeof(float) * N));
    FB_HIP(hipMemset(bgAcc, 0, sizeof(float)));
    FB_HIP(hipMemset(ncoll, 0, sizeof(int32_t)));
    FB_HIP(hipMemset(err, 0, sizeof(int32_t)));
    FB_HIP(hipMemset(counters, 0, sizeof(unsigned long long) * 4));
    b->steps.assign((size_t)B, INT32_MAX);
    FB_HIP(hipMemcpy(stepsD, b->steps.data(), sizeof(int32_t) * B, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(kb_init, dim3(1024), dim3(256), 0, b->stream, d);
    FB_HIP(hipGetLastError());
    FB_HIP(hipStreamSynchronize(b->stream));
    b->typeOf.assign(typeOf, typeOf + N);
    *out = b;
    return FB_OK;
}

extern "C" fb_status fb_batch_destroy(fb_batch *b) {
    FB_REQUIRE_B(b, FB_ERR_ARG, "fb_batch_destroy: null batch");
    (void)hipStreamSynchronize(b->stream);
    batch_la_release((const void *)b->d.v);
    for (void *p : b->allocs) (void)hipFree(p);
    void *rec[] = {b->recCnt, b->recIds, b->recMasks, b->recF};
    for (void *p : rec)
        if (p) (void)hipFree(p);
    (void)hipStreamDestroy(b->ownStream);
    fb_graph_release(b->g);
    delete b;
    return FB_OK;
}

extern "C" fb_status fb_batch_set_drive(fb_batch *b, const int32_t *idx, int32_t n, double hz) {
    FB_REQUIRE_B(b && n >= 0 && (n == 0 || idx), FB_ERR_ARG, "fb_batch_set_drive: bad arguments");
    std::vector<float> drive((size_t)b->d.N);
    FB_HIP(hipMemcpy(drive.data(), b->d.drive, sizeof(float) * b->d.N, hipMemcpyDeviceToHost));
    for (int32_t q = 0; q < n; q++) {
        const int32_t i = idx[q];
        FB_REQUIRE_B(i >= 0 && i < b->d.N, FB_ERR_ARG, "fb_batch_set_drive: index %d out of range", i);
        drive[i] = (float)hz;
        auto it = std::find(b->drivenList.begin(), b->drivenList.end(), i);
        if (hz > 0) { if (it == b->drivenList.end()) b->drivenList.push_back(i); }
        else if (it != b->drivenList.end()) b->drivenList.erase(it);
    }
    FB_REQUIRE_B(b->drivenList.size() <= 2047, FB_ERR_UNSUPPORTED, "fb_batch_set_drive: at most 2047 driven neurons");
    FB_HIP(hipMemcpy(b->d.drive, drive.data(), sizeof(float) * b->d.N, hipMemcpyHostToDevice));
    std::vector<int32_t> pos((size_t)b->d.N, -1);
    for (size_t q = 0; q < b->drivenList.size(); q++) pos[b->drivenList[q]] = (int32_t)q;
    FB_HIP(hipMemcpy((void *)b->d.driven, b->drivenList.data(), sizeof(int32_t) * std::max<size_t>(1, b->drivenList.size()),
                     hipMemcpyHostToDevice));
    FB_HIP(hipMemcpy((void *)b->d.drvPos, pos.data(), sizeof(int32_t) * b->d.N, hipMemcpyHostToDevice));
    b->d.nDriven = (int32_t)b->drivenList.size();
    b->drivenVersion++;
    return FB_OK;
}

// per-member thr / bias (mV) for a list of neurons
extern "C" fb_status fb_batch_set_member(fb_batch *b, int32_t which, int32_t member, const int32_t *idx, int32_t n,
                                         double value) {
    FB_REQUIRE_B(b && member >= 0 && member < b->d.B && n >= 0 && (n == 0 || idx), FB_ERR_ARG,
                 "fb_batch_set_member: bad arguments");
    FB_REQUIRE_B(which == FB_THR || which == FB_BIAS, FB_ERR_ARG, "fb_batch_set_member: FB_THR or FB_BIAS only");
    for (int32_t q = 0; q < n; q++)
        FB_REQUIRE_B(idx[q] >= 0 && idx[q] < b->d.N, FB_ERR_ARG, "fb_batch_set_member: index %d out of range", idx[q]);
    if (n == 0) return FB_OK;
    FB_HIP(hipMemcpyAsync(b->idxScratch, idx, sizeof(int32_t) * n, hipMemcpyHostToDevice, b->stream));
    hipLaunchKernelGGL(kb_set_member, dim3((n + 255) / 256), dim3(256), 0, b->stream,
                       which == FB_THR ? b->d.thr : b->d.bias, b->d.B, member, (const int32_t *)b->idxScratch, n,
                       (float)value);
    FB_HIP(hipGetLastError());
    FB_HIP(hipStreamSynchronize(b->stream));
    return FB_OK;
}

extern "C" fb_status fb_batch_set_steps(fb_batch *b, const int32_t *steps) {
    FB_REQUIRE_B(b && steps, FB_ERR_ARG, "fb_batch_set_steps: null argument");
    b->steps.assign(steps, steps + b->d.B);
    // run lengths count from the current step
    std::vector<int32_t> abs((size_t)b->d.B);
    for (int32_t q = 0; q < b->d.B; q++) abs[q] = steps[q] < 0 ? INT32_MAX : b->t + steps[q];
    FB_HIP(hipMemcpy(b->d.steps, abs.data(), sizeof(int32_t) * b->d.B, hipMemcpyHostToDevice));
    return FB_OK;
}

static fb_status batch_step(fb_batch *b) {
    BatchDev &d = b->d;
    const int32_t head = b->head, slot = (head + d.nslots - 1) % d.nslots;
    hipStream_t st = b->stream;
    hipLaunchKernelGGL(kb_deliver, dim3(256), dim3(256), 0, st, d, head, b->t);
    d.multiDrv = !d.liveBg && d.k.bgEvents == 0.f && d.nDriven > 256 && d.drvPart ? 1 : 0;
    if (d.multiDrv) {
        hipLaunchKernelGGL(kb_driven_count, dim3(DRV_WG), dim3(1024), 0, st, d, b->t);
        hipLaunchKernelGGL(kb_driven_multi, dim3(DRV_WG), dim3(1024), 0, st, d, slot, b->t);
    } else {
        hipLaunchKernelGGL(kb_driven, dim3(1), dim3(1024), 0, st, d, slot, b->t);
    }
    hipLaunchKernelGGL(kb_membrane, dim3((d.N + 3) / 4), dim3(256), 0, st, d, slot, head, b->t);
    if (d.poolOf) hipLaunchKernelGGL(kb_pool_loop, dim3(1), dim3(256), 0, st, d, slot);
    FB_HIP(hipGetLastError());
    b->lastSlot = slot;
    b->head = (head + 1) % d.nslots;
    b->t++;
    return FB_OK;
}

// Runs nsteps; if ids/masks are given, records every step's active neurons with
// their spike and forced masks (per step: count into nactive[step], entries
// appended at most cap in total; the returned total may exceed cap).
// nsteps on the per-step kernels; with recording, appends every step's
// active neurons and masks after *tot entries
static fb_status run_per_step(fb_batch *b, int32_t nsteps, int32_t *nactive, int32_t *ids, uint64_t *masks,
                              uint64_t *fmasks, int64_t cap, int64_t *tot) {
    const bool rec = nactive != nullptr;
    if (rec && !b->recCnt) {
        FB_HIP(hipMalloc(&b->recCnt, sizeof(int32_t)));
        FB_HIP(hipMalloc(&b->recIds, sizeof(int32_t) * b->d.N));
        FB_HIP(hipMalloc(&b->recMasks, sizeof(uint64_t) * b->d.N));
        FB_HIP(hipMalloc(&b->recF, sizeof(uint64_t) * b->d.N));
    }
    for (int32_t s = 0; s < nsteps; s++) {
        fb_status st = batch_step(b);
        if (st) return st;
        if (rec) {
            hipLaunchKernelGGL(kb_record, dim3(16), dim3(256), 0, b->stream, b->d, b->lastSlot, b->recCnt, b->recIds,
                               b->recMasks, b->recF, b->d.N);
            int32_t n = 0;
            FB_HIP(hipMemcpyAsync(&n, b->recCnt, sizeof(int32_t), hipMemcpyDeviceToHost, b->stream));
            FB_HIP(hipStreamSynchronize(b->stream));
            nactive[s] = n;
            const int64_t take = std::max<int64_t>(0, std::min<int64_t>(n, cap - *tot));
            if (take > 0 && ids) {
                FB_HIP(hipMemcpyAsync(ids + *tot, b->recIds, sizeof(int32_t) * take, hipMemcpyDeviceToHost, b->stream));
                FB_HIP(hipMemcpyAsync(masks + *tot, b->recMasks, sizeof(uint64_t) * take, hipMemcpyDeviceToHost,
                                      b->stream));
                FB_HIP(hipMemcpyAsync(fmasks + *tot, b->recF, sizeof(uint64_t) * take, hipMemcpyDeviceToHost,
                                      b->stream));
                FB_HIP(hipStreamSynchronize(b->stream));
            }
            *tot += n;
        }
    }
    return FB_OK;
}

extern "C" fb_status fb_batch_set_lanes(fb_batch *b, int32_t signGain, const float *laneDrive) {
    FB_REQUIRE_B(b, FB_ERR_ARG, "fb_batch_set_lanes: null ensemble");
    BatchDev &d = b->d;
    // the pool loop writes the per-lane drive after every step: it needs the buffer
    FB_REQUIRE_B(laneDrive || !d.poolOf, FB_ERR_ARG,
                 "fb_batch_set_lanes: the pool loop needs per-lane drive; clear it first (fb_batch_set_pool_feedback "
                 "with no pools)");
    FB_HIP(hipStreamSynchronize(b->stream));
    d.signGain = signGain ? 1 : 0;
    if (laneDrive) {
        const int32_t N = d.N, B = d.B;
        if (!b->laneDriveBuf) {
            fb_status st = balloc(b, &b->laneDriveBuf, (size_t)N * B);
            if (st) return st;
        }
        d.laneDrive = b->laneDriveBuf;
        std::vector<float> t((size_t)N * B);                 // [B][N] -> [N][B]
        for (int32_t q = 0; q < B; q++)
            for (int32_t i = 0; i < N; i++) t[(size_t)i * B + q] = laneDrive[(size_t)q * N + i];
        FB_HIP(hipMemcpy(d.laneDrive, t.data(), sizeof(float) * t.size(), hipMemcpyHostToDevice));
    } else {
        d.laneDrive = nullptr;          // the buffer stays in laneDriveBuf
    }
    return FB_OK;
}

extern "C" fb_status fb_batch_size(const fb_batch *b, int32_t *N, int32_t *B, int32_t *nTypes) {
    FB_REQUIRE_B(b && N && B && nTypes, FB_ERR_ARG, "fb_batch_size: bad arguments");
    *N = b->d.N;
    *B = b->d.B;
    *nTypes = b->d.nTypes;
    return FB_OK;
}

extern "C" fb_status fb_batch_set_rng(fb_batch *b, const uint32_t *states) {
    FB_REQUIRE_B(b && states, FB_ERR_ARG, "fb_batch_set_rng: bad arguments");
    for (int32_t q = 0; q < b->d.B; q++)
        FB_REQUIRE_B(states[q] != 0, FB_ERR_ARG, "fb_batch_set_rng: member %d has a zero state", q);
    FB_HIP(hipStreamSynchronize(b->stream));
    FB_HIP(hipMemcpy(b->d.rng, states, sizeof(uint32_t) * b->d.B, hipMemcpyHostToDevice));
    return FB_OK;
}

extern "C" fb_status fb_batch_set_pool_feedback(fb_batch *b, const uint8_t *poolOf, int32_t nPools,
                                                const int32_t *targets, int32_t nTargets, float base, float gain,
                                                const int32_t *delay, int32_t recordCap) {
    FB_REQUIRE_B(b, FB_ERR_ARG, "fb_batch_set_pool_feedback: null ensemble");
    BatchDev &d = b->d;
    FB_HIP(hipStreamSynchronize(b->stream));
    if (d.poolOf) {                    // the previous loop's buffers
        for (const void *p : {(const void *)d.poolOf, (const void *)d.loopTargets, (const void *)d.loopDelay,
                              (const void *)d.loopState, (const void *)d.loopHist, (const void *)d.loopRec,
                              (const void *)d.loopCells}) {
            auto it = std::find(b->allocs.begin(), b->allocs.end(), (void *)p);
            if (it != b->allocs.end()) { b->allocs.erase(it); (void)hipFree((void *)p); }
        }
        d.poolOf = nullptr; d.loopTargets = d.loopDelay = d.loopCells = nullptr;
        d.loopState = nullptr; d.loopHist = nullptr; d.loopRec = nullptr;
    }
    if (!poolOf) return FB_OK;
    FB_REQUIRE_B(d.laneDrive, FB_ERR_ARG, "fb_batch_set_pool_feedback: set per-lane drive first (fb_batch_set_lanes)");
    FB_REQUIRE_B(nPools >= 1 && nPools <= 8 && nTargets >= 0 && (nTargets == 0 || targets) && delay && recordCap >= 0,
                 FB_ERR_ARG, "fb_batch_set_pool_feedback: bad arguments");
    int32_t maxDelay = 1;
    for (int32_t q = 0; q < d.B; q++) {
        FB_REQUIRE_B(delay[q] >= 1, FB_ERR_ARG, "fb_batch_set_pool_feedback: member %d has delay < 1", q);
        maxDelay = std::max(maxDelay, delay[q]);
    }
    for (int32_t q = 0; q < nTargets; q++)
        FB_REQUIRE_B(targets[q] >= 0 && targets[q] < d.N, FB_ERR_ARG, "fb_batch_set_pool_feedback: target %d out of range", q);
    const int32_t ring = maxDelay + 1;
    std::vector<int32_t> cells;
    for (int32_t i = 0; i < d.N; i++)
        if ((int32_t)poolOf[i] < nPools) cells.push_back(i);
    int32_t *cl;
    uint8_t *pool;
    int32_t *tg, *dl, *state;
    uint32_t *hist;
    uint16_t *rec;
    fb_status st;
    if ((st = balloc(b, &pool, (size_t)d.N)) || (st = balloc(b, &tg, (size_t)std::max(nTargets, 1))) ||
        (st = balloc(b, &dl, (size_t)d.B)) || (st = balloc(b, &state, 2)) || (st = balloc(b, &hist, (size_t)d.B * ring)) ||
        (st = balloc(b, &rec, (size_t)d.B * std::max(recordCap, 1) * nPools)) ||
        (st = balloc(b, &cl, std::max<size_t>(cells.size(), 1))))
        return st;
    if (!cells.empty()) FB_HIP(hipMemcpy(cl, cells.data(), sizeof(int32_t) * cells.size(), hipMemcpyHostToDevice));
    FB_HIP(hipMemcpy(pool, poolOf, (size_t)d.N, hipMemcpyHostToDevice));
    if (nTargets) FB_HIP(hipMemcpy(tg, targets, sizeof(int32_t) * nTargets, hipMemcpyHostToDevice));
    FB_HIP(hipMemcpy(dl, delay, sizeof(int32_t) * d.B, hipMemcpyHostToDevice));
    FB_HIP(hipMemset(state, 0, sizeof(int32_t) * 2));
    FB_HIP(hipMemset(hist, 0, sizeof(uint32_t) * (size_t)d.B * ring));
    d.poolOf = pool; d.loopTargets = tg; d.loopDelay = dl; d.loopState = state; d.loopHist = hist; d.loopRec = rec;
    d.loopPools = nPools; d.loopNTargets = nTargets; d.loopRing = ring; d.loopRecCap = recordCap;
    d.loopBase = base; d.loopGain = gain; d.loopCells = cl; d.loopNCells = (int32_t)cells.size();
    return FB_OK;
}

extern "C" fb_status fb_batch_pool_counts(fb_batch *b, uint16_t *dst, int32_t cap, int32_t *nRecords) {
    FB_REQUIRE_B(b && nRecords && (cap == 0 || dst), FB_ERR_ARG, "fb_batch_pool_counts: bad arguments");
    const BatchDev &d = b->d;
    FB_REQUIRE_B(d.poolOf, FB_ERR_ARG, "fb_batch_pool_counts: no pool loop is set");
    int32_t state[2];
    FB_HIP(hipMemcpyAsync(state, d.loopState, sizeof state, hipMemcpyDeviceToHost, b->stream));
    FB_HIP(hipStreamSynchronize(b->stream));
    *nRecords = state[1];
    const int32_t n = std::min(cap, state[1]);
    for (int32_t q = 0; q < d.B && n > 0; q++)
        FB_HIP(hipMemcpy(dst + (size_t)q * n * d.loopPools, d.loopRec + (size_t)q * d.loopRecCap * d.loopPools,
                         sizeof(uint16_t) * (size_t)n * d.loopPools, hipMemcpyDeviceToHost));
    return FB_OK;
}

// Runs nsteps; if ids/masks are given, records every step's active neurons with
// their spike and forced masks (per step: count into nactive[step], entries
// appended at most cap in total; the returned total may exceed cap).
// the autotuner's choices, by (graph, neurons, members, driven neurons)
struct TuneKey {
    const void *graph;
    int32_t N, B, nDriven;
    bool operator<(const TuneKey &o) const {
        return std::tie(graph, N, B, nDriven) < std::tie(o.graph, o.N, o.B, o.nDriven);
    }
};
static std::map<TuneKey, int32_t> g_tuned;

extern "C" fb_status fb_batch_run(fb_batch *b, int32_t nsteps, int32_t *nactive, int32_t *ids, uint64_t *masks,
                                  uint64_t *fmasks, int64_t cap, int64_t *total) {
    FB_REQUIRE_B(b && nsteps >= 0, FB_ERR_ARG, "fb_batch_run: bad arguments");
    BatchDev &d = b->d;
    const bool rec = nactive != nullptr;
    const int32_t D = d.nslots - 1;
    std::string why;
    const bool ok = batch_la_supported(d, &why);
    bool la = b->engine == FB_ENGINE_PERSISTENT || (b->engine == FB_ENGINE_AUTO && ok && nsteps >= 2 * D);
    if (la && !ok) {
        set_error("fb_batch_run: the lookahead engine does not take this ensemble: %s", why.c_str());
        return FB_ERR_UNSUPPORTED;
    }
    int64_t tot = 0;
    int32_t s0 = 0;
    fb_status st;
    // with recording, launches are cut to whole epochs of at most ~512 steps
    const int32_t recEpochs = std::max(1, 512 / std::max(1, D));
    uint64_t *recM = nullptr, *recF = nullptr;
    std::vector<uint64_t> hM, hF;
    // The autotuner: the first run long enough runs its first epochs on every
    // shape that fits, twice round, keeps the fastest for the rest of the run
    // and every later one, and leaves the choice for later ensembles of the
    // same graph, size and driven set.  All shapes give the same bits, so the
    // trials are the run's own epochs.
    constexpr int32_t TRIAL_EPOCHS = 2;
    std::vector<int32_t> fits;
    for (int k = 0; k < LA_NSHAPES; k++)
        if (la && batch_la_shape_fits(d, k)) fits.push_back(k);
    const TuneKey tkey{(const void *)d.g.edges, d.N, d.B, d.nDriven};
    if (la && b->laForced < 0 && b->laShape < 0 && b->tunePlan.empty() && !fits.empty()) {
        auto hit = g_tuned.find(tkey);
        if (hit != g_tuned.end()) {
            b->laShape = hit->second;
        } else if (fits.size() == 1) {
            b->laShape = fits[0];
        } else if ((nsteps - s0) / D >= 2 + 2 * (int32_t)fits.size() * TRIAL_EPOCHS) {
            b->tunePlan.push_back({fits[0], 1, false});        // out of the transient of a fresh start
            for (int round = 0; round < 2; round++)
                for (int32_t k : fits) b->tunePlan.push_back({k, TRIAL_EPOCHS, true});
            for (int k = 0; k < LA_NSHAPES; k++) { b->tuneMs[k] = 0.0; b->tuneEpochs[k] = 0; }
        }
    }
    while (la && nsteps - s0 >= D) {
        int32_t epochs = (nsteps - s0) / D;
        int32_t shape = b->laForced >= 0 ? b->laForced : b->laShape >= 0 ? b->laShape : (fits.empty() ? 0 : fits[0]);
        bool timed = false;
        const bool trial = b->laForced < 0 && !b->tunePlan.empty();
        if (trial) {
            const fb_batch::Trial tr = b->tunePlan.front();
            b->tunePlan.erase(b->tunePlan.begin());
            shape = tr.shape;
            epochs = std::min(epochs, tr.epochs);
            timed = tr.timed;
        }
        if (rec) {
            epochs = std::min(epochs, recEpochs);
            const size_t n = (size_t)epochs * D * d.N;
            if (!recM) {
                FB_HIP(hipMalloc(&recM, sizeof(uint64_t) * (size_t)recEpochs * D * d.N));
                FB_HIP(hipMalloc(&recF, sizeof(uint64_t) * (size_t)recEpochs * D * d.N));
            }
            FB_HIP(hipMemsetAsync(recM, 0, sizeof(uint64_t) * n, b->stream));
            FB_HIP(hipMemsetAsync(recF, 0, sizeof(uint64_t) * n, b->stream));
        }
        int32_t done = 0;
        const double ms0 = b->laMs;
        st = batch_la_launch(d, shape, b->t, b->head, epochs, b->drivenVersion, recM, recF, b->stream, &done,
                             &b->laMs, b->laPhaseUs);
        if (st == FB_ERR_UNSUPPORTED && b->engine == FB_ENGINE_AUTO) {
            // the device refused the cooperative launch: the per-step kernels
            la = false;
            break;
        }
        if (st) return st;
        b->laLaunches++;
        if (timed && done == epochs) {
            b->tuneMs[shape] += b->laMs - ms0;
            b->tuneEpochs[shape] += done;
        }
        if (trial && b->tunePlan.empty()) {
            int best = -1;
            for (int32_t k : fits)
                if (b->tuneEpochs[k] > 0 &&
                    (best < 0 || b->tuneMs[k] / b->tuneEpochs[k] < b->tuneMs[best] / b->tuneEpochs[best]))
                    best = k;
            b->laShape = best >= 0 ? best : fits[0];
            g_tuned[tkey] = b->laShape;
        }
        const int32_t m = done * D;
        if (rec && m > 0) {
            hM.resize((size_t)m * d.N);
            hF.resize((size_t)m * d.N);
            FB_HIP(hipMemcpy(hM.data(), recM, sizeof(uint64_t) * hM.size(), hipMemcpyDeviceToHost));
            FB_HIP(hipMemcpy(hF.data(), recF, sizeof(uint64_t) * hF.size(), hipMemcpyDeviceToHost));
            for (int32_t q = 0; q < m; q++) {
                int32_t cnt = 0;
                for (int32_t i = 0; i < d.N; i++) {
                    const uint64_t mk = hM[(size_t)q * d.N + i];
                    if (!mk) continue;
                    if (tot < cap && ids) {
                        ids[tot] = i;
                        masks[tot] = mk;
                        fmasks[tot] = hF[(size_t)q * d.N + i];
                    }
                    tot++;
                    cnt++;
                }
                nactive[s0 + q] = cnt;
            }
        }
        b->t += m;
        b->head = (b->head + m) % d.nslots;
        b->lastSlot = (b->head + d.nslots - 1) % d.nslots;
        b->laSteps += m;
        s0 += m;
        if (done < epochs) {
            // a bucket overflowed: this epoch on the per-step kernels
            const int32_t k = std::min(D, nsteps - s0);
            if ((st = run_per_step(b, k, rec ? nactive + s0 : nullptr, ids, masks, fmasks, cap, &tot))) return st;
            b->laFallback += k;
            s0 += k;
        }
    }
    if (recM) {
        (void)hipFree(recM);
        (void)hipFree(recF);
    }
    if ((st = run_per_step(b, nsteps - s0, rec ? nactive + s0 : nullptr, ids, masks, fmasks, cap, &tot))) return st;
    if (total) *total = tot;
    int32_t err = 0;
    FB_HIP(hipMemcpyAsync(&err, b->d.err, sizeof(int32_t), hipMemcpyDeviceToHost, b->stream));
    FB_HIP(hipStreamSynchronize(b->stream));
    FB_REQUIRE_B(!(err & FB_DEVERR_COLL_OVERFLOW), FB_ERR_OVERFLOW, "fb_batch_run: arrival chain pool overflowed");
    FB_REQUIRE_B(!(err & 2), FB_ERR_UNSUPPORTED, "fb_batch_run: more than 2047 draws in one step");
    return FB_OK;
}

// The lookahead engine's regular-block shape: 72, 80, 96 or 128 neurons per
// block, or 0 for the autotuner's choice (measured again on the next run long
// enough).  All shapes give the same bits.
extern "C" fb_status fb_batch_set_shape(fb_batch *b, int32_t neuronsPerBlock) {
    FB_REQUIRE_B(b, FB_ERR_ARG, "fb_batch_set_shape: null batch");
    b->tunePlan.clear();
    if (neuronsPerBlock == 0) {
        b->laForced = -1;
        b->laShape = -1;
        g_tuned.erase(TuneKey{(const void *)b->d.g.edges, b->d.N, b->d.B, b->d.nDriven});
        return FB_OK;
    }
    for (int k = 0; k < LA_NSHAPES; k++)
        if (batch_la_shape_neurons(k) == neuronsPerBlock) {
            FB_REQUIRE_B(batch_la_shape_fits(b->d, k), FB_ERR_UNSUPPORTED,
                         "fb_batch_set_shape: %d neurons per block does not fit the device for this ensemble",
                         neuronsPerBlock);
            b->laForced = k;
            return FB_OK;
        }
    set_error("fb_batch_set_shape: %d is not a shape (72, 80, 96, 128 or 0)", neuronsPerBlock);
    return FB_ERR_ARG;
}

// FB_ENGINE_AUTO (lookahead when the ensemble allows it), FB_ENGINE_LAUNCH
// (per-step kernels) or FB_ENGINE_PERSISTENT (lookahead, or an error)
extern "C" fb_status fb_batch_set_engine(fb_batch *b, int32_t engine) {
    FB_REQUIRE_B(b, FB_ERR_ARG, "fb_batch_set_engine: null batch");
    FB_REQUIRE_B(engine == FB_ENGINE_AUTO || engine == FB_ENGINE_LAUNCH || engine == FB_ENGINE_PERSISTENT, FB_ERR_ARG,
                 "fb_batch_set_engine: AUTO, LAUNCH or PERSISTENT");
    b->engine = engine;
    return FB_OK;
}

// "lookahead_launches", "lookahead_kernel_ms", "lookahead_steps",
// "fallback_steps", "lookahead_supported"
extern "C" fb_status fb_batch_stat(fb_batch *b, const char *name, double *value) {
    FB_REQUIRE_B(b && name && value, FB_ERR_ARG, "fb_batch_stat: null argument");
    const std::string n(name);
    if (n == "lookahead_launches") *value = (double)b->laLaunches;
    else if (n == "lookahead_kernel_ms") *value = b->laMs;
    else if (n == "lookahead_steps") *value = (double)b->laSteps;
    else if (n == "fallback_steps") *value = (double)b->laFallback;
    else if (n == "lookahead_supported") *value = batch_la_supported(b->d, nullptr) ? 1.0 : 0.0;
    else if (n == "pools") *value = b->d.poolOf ? (double)b->d.loopPools : 0.0;
    else if (n == "lookahead_shape") {
        const int k = b->laForced >= 0 ? b->laForced : b->laShape;
        *value = k >= 0 ? (double)batch_la_shape_neurons(k) : 0.0;
    } else if (n.rfind("la_phase_us_", 0) == 0) {
        // la_phase_us_<drv|reg>_<prologue|gather|sim|push|barrier>: the first
        // driven and the first regular workgroup of the ensemble lookahead kernel
        static const char *blk[2] = {"drv_", "reg_"};
        static const char *ph[5] = {"prologue", "gather", "sim", "push", "barrier"};
        int idx = -1;
        for (int q = 0; q < 2; q++)
            for (int k = 0; k < 5; k++)
                if (n == std::string("la_phase_us_") + blk[q] + ph[k]) idx = q * 5 + k;
        if (idx < 0) { set_error("fb_batch_stat: unknown statistic '%s'", name); return FB_ERR_ARG; }
        *value = b->laPhaseUs[idx];
    } else if (n.rfind("autotune_us_per_epoch_", 0) == 0) {
        // the autotuner's measurement of one shape (0: not measured)
        const int32_t nb = std::atoi(n.c_str() + 22);
        *value = 0.0;
        for (int k = 0; k < LA_NSHAPES; k++)
            if (batch_la_shape_neurons(k) == nb && b->tuneEpochs[k] > 0) *value = 1e3 * b->tuneMs[k] / b->tuneEpochs[k];
    } else {
        set_error("fb_batch_stat: unknown statistic '%s'", name);
        return FB_ERR_ARG;
    }
    return FB_OK;
}

// one member's state array (FB_V .. FB_THR, FB_SPIKECOUNT)
extern "C" fb_status fb_batch_get_array(fb_batch *b, int32_t which, int32_t member, void *dst) {
    FB_REQUIRE_B(b && dst && member >= 0 && member < b->d.B, FB_ERR_ARG, "fb_batch_get_array: bad arguments");
    float *src = nullptr;
    switch (which) {
        case FB_V: src = b->d.v; break;
        case FB_GE: src = b->d.gE; break;
        case FB_GI: src = b->d.gI; break;
        case FB_REFR: src = b->d.refr; break;
        case FB_TRACE: src = b->d.trace; break;
        case FB_ADAPT: src = b->d.adapt; break;
        case FB_RES: src = b->d.res; break;
        case FB_BIAS: src = b->d.bias; break;
        case FB_THR: src = b->d.thr; break;
        case FB_SPIKECOUNT: src = (float *)b->d.spikeCount; break;
        default: set_error("fb_batch_get_array: unknown array %d", which); return FB_ERR_ARG;
    }
    FB_HIP(hipStreamSynchronize(b->stream));
    // strided copy: one column of the [N][B] layout
    FB_HIP(hipMemcpy2D(dst, sizeof(float), src + member, sizeof(float) * b->d.B, sizeof(float), b->d.N,
                       hipMemcpyDeviceToHost));
    return FB_OK;
}

extern "C" fb_status fb_batch_rng(fb_batch *b, uint32_t *rng) {
    FB_REQUIRE_B(b && rng, FB_ERR_ARG, "fb_batch_rng: null argument");
    FB_HIP(hipStreamSynchronize(b->stream));
    FB_HIP(hipMemcpy(rng, b->d.rng, sizeof(uint32_t) * b->d.B, hipMemcpyDeviceToHost));
    return FB_OK;
}

extern "C" fb_status fb_batch_counters(fb_batch *b, int64_t *edge_events, int64_t *synapse_events, int64_t *spikes,
                                       int64_t *steps) {
    FB_REQUIRE_B(b, FB_ERR_ARG, "fb_batch_counters: null batch");
    unsigned long long c[4];
    FB_HIP(hipStreamSynchronize(b->stream));
    FB_HIP(hipMemcpy(c, b->d.counters, sizeof c, hipMemcpyDeviceToHost));
    if (edge_events) *edge_events = (int64_t)c[0];
    if (synapse_events) *synapse_events = (int64_t)c[1];
    if (spikes) *spikes = (int64_t)c[2];
    if (steps) *steps = b->t;
    return FB_OK;
}

extern "C" fb_status fb_batch_set_stream(fb_batch *b, void *stream) {
    FB_REQUIRE_B(b, FB_ERR_ARG, "fb_batch_set_stream: null batch");
    FB_HIP(hipStreamSynchronize(b->stream));
    b->stream = stream ? (hipStream_t)stream : b->ownStream;
    return FB_OK;
}
