// fb_step.hip -- the per-step kernel sequence of the forward LIF engine.
//
// One step of the reference's lif_step (src/wasm/lif.c:29-80) runs as
//   k_deliver   arrivals of ring[head] scatter into per-(sign, target)
//               accumulators; the first arrival on a target keeps its exact
//               addend f32(W*s) and order key, later ones go to the target's
//               own run (one entry per in-edge).  No floating-point atomics:
//               the sum is formed later, in arrival order, so it is
//               bit-identical to lif.c's serial loop.  Zero-weight edges are
//               skipped where that is exact (DevBrain::skipZeroW).
//   k_driven    one workgroup: background draws and forced spikes.  lif.c
//               draws from one serial xorshift32 stream; each thread counts the
//               non-refractory driven neurons of its slice, a workgroup scan
//               gives each slice its position in the stream, and the stream is
//               produced 2048 states per round by GF(2) leapfrog matrices.
//   k_deliver_driven  both in one launch when there is no depression, no
//               background and at most 4096 driven neurons: a flat list of
//               edge events over 128 workgroups, and one workgroup whose
//               threads jump straight to their first draw.
//   k_long      one workgroup per (class, target) with more than 16
//               arrivals: ranks them by order key and adds them in order.
//   k_membrane  dense pass: apply the step's arrivals (single arrival: one
//               add; several: the ordered sum), background hits, the membrane
//               and decay update, the threshold test, and wave-ballot
//               compaction of the spikes into a per-block staging list.
//   k_compact   block prefix over the staged counts (only blocks that staged
//               spikes, and the last block, compute one): the ring slot gets
//               the forced spikes first, then threshold spikes in ascending
//               index, which is lif.c's output order.
#include "fb_internal.h"

namespace fb {

// ------------------------------------------------------------------ deliver

// One arrival's edge event into the accumulators: the first on its (class,
// target) keeps addend and key there, later ones go to the target's run.
__device__ __forceinline__ void deliver_edge(const DevBrain &d, uint2 ed, float s, uint64_t key) {
    // a +0 weight adds -0 to gI, which changes nothing, and +0 to gE, which
    // changes nothing while no excitatory conductance can be -0
    if (ed.y == 0u && (s < 0.f || d.skipZeroW)) return;
    const float addend = __uint_as_float(ed.y) * s;
    const int32_t ci = (s > 0 ? 0 : d.N) + (int32_t)ed.x;
    const int32_t slot = atomicAdd(&d.cnt[ci], 1);
    if (slot == LONG_RUN) d.longList[atomicAdd(&d.sc->nLong, 1)] = ci;
    if (slot == 0) {
        d.add[ci] = addend;
        d.key[ci] = key;
    } else {
        // each in-edge of the target arrives at most once per step
        const uint32_t r0 = d.g.inOff[ed.x], deg = d.g.inOff[ed.x + 1] - r0;
        if ((uint32_t)(slot - 1) < deg) {
            const int64_t at = (s > 0 ? 0 : d.g.E) + r0 + (slot - 1);
            d.runKey[at] = key;
            d.runAdd[at] = addend;
        } else {
            atomicOr(&d.sc->err, FB_DEVERR_RUN_OVERFLOW);
        }
    }
}

// Arrivals row by row: a workgroup per arriving neuron, which also applies
// its synaptic depression once every thread has read res.
__global__ void __launch_bounds__(256) k_deliver(DevBrain d, int32_t head) {
    const int32_t A = d.ringCount[head];
    const int32_t *list = d.ring + (int64_t)head * d.ringStride;
    for (int32_t k = blockIdx.x; k < A; k += gridDim.x) {
        const int32_t pre = list[k];
        const float s = d.g.sign[pre] * d.res[pre];
        if (s == 0.f) continue;                       // uniform across the block
        const uint32_t a = d.g.indptr[pre], e = d.g.indptr[pre + 1];
        __syncthreads();                              // every thread has read res[pre]
        if (threadIdx.x == 0) {
            d.res[pre] = d.res[pre] - d.k.depU * d.res[pre];
            atomicAdd(&d.sc->edgeEvents, (unsigned long long)(e - a));
            atomicAdd(&d.sc->synEvents, (unsigned long long)d.g.rowSyn[pre]);
        }
        for (uint32_t j = a + threadIdx.x; j < e; j += blockDim.x)
            deliver_edge(d, d.g.edges[j], s, ((uint64_t)(uint32_t)k << 32) | (uint64_t)(j - a));
    }
}

// Arrivals as one flat list of edge events (without synaptic depression,
// when res never changes on arrival): every workgroup builds the table of the
// step's rows in LDS (start, length prefix, sign x res), and thread g of the
// grid takes edge events g, g + grid, ... , so long rows spread over the
// whole chip.  Up to DF_CAP arriving neurons; more go row by row.
constexpr int DF_T = 1024;
constexpr int DF_CAP = 2048;

// blocks 0 .. nblk-1 of the grid share the flat list
__device__ __forceinline__ void deliver_flat(const DevBrain &d, int32_t head, int32_t nblk) {
    __shared__ int32_t rowOff[DF_CAP + 1];
    __shared__ uint32_t rowA[DF_CAP];
    __shared__ float rowS[DF_CAP];
    __shared__ int32_t scan[64 + DF_T / 64];
    constexpr int RPT = DF_CAP / DF_T;
    const int tid = threadIdx.x;
    const int32_t A = d.ringCount[head];
    const int32_t *list = d.ring + (int64_t)head * d.ringStride;
    if (A > DF_CAP) {
        for (int32_t k = blockIdx.x; k < A; k += nblk) {
            const int32_t pre = list[k];
            const float s = d.g.sign[pre] * d.res[pre];
            if (s == 0.f) continue;
            const uint32_t a = d.g.indptr[pre], e = d.g.indptr[pre + 1];
            if (threadIdx.x == 0) {
                atomicAdd(&d.sc->edgeEvents, (unsigned long long)(e - a));
                atomicAdd(&d.sc->synEvents, (unsigned long long)d.g.rowSyn[pre]);
            }
            for (uint32_t j = a + threadIdx.x; j < e; j += blockDim.x)
                deliver_edge(d, d.g.edges[j], s, ((uint64_t)(uint32_t)k << 32) | (uint64_t)(j - a));
        }
        return;
    }
    int32_t len[RPT];
    int32_t mySum = 0;
    unsigned long long syn = 0;
#pragma unroll
    for (int v = 0; v < RPT; v++) {
        const int32_t r = RPT * tid + v;
        len[v] = 0;
        if (r < A) {
            const int32_t pre = list[r];
            const float s = d.g.sign[pre] * d.res[pre];
            const uint32_t a = d.g.indptr[pre];
            if (s != 0.f) {
                len[v] = (int32_t)(d.g.indptr[pre + 1] - a);
                syn += d.g.rowSyn[pre];
            }
            rowA[r] = a;
            rowS[r] = s;
        }
        mySum += len[v];
    }
    int32_t total;
    int32_t off = wg_exclusive_scan<DF_T>(mySum, scan, &total);
#pragma unroll
    for (int v = 0; v < RPT; v++) {
        const int32_t r = RPT * tid + v;
        if (r < A) rowOff[r] = off;
        off += len[v];
    }
    if (tid == 0) rowOff[A] = total;
    if (blockIdx.x == 0 && mySum) {
        atomicAdd(&d.sc->edgeEvents, (unsigned long long)mySum);
        atomicAdd(&d.sc->synEvents, syn);
    }
    __syncthreads();
    for (int32_t g = blockIdx.x * DF_T + tid; g < total; g += nblk * DF_T) {
        int32_t lo = 0, hi = A;   // last row with rowOff <= g (rows of length 0 share offsets)
        while (hi - lo > 1) {
            const int32_t mid = (lo + hi) >> 1;
            if (rowOff[mid] <= g) lo = mid; else hi = mid;
        }
        const uint32_t pos = (uint32_t)(g - rowOff[lo]);
        deliver_edge(d, d.g.edges[rowA[lo] + pos], rowS[lo], ((uint64_t)(uint32_t)lo << 32) | (uint64_t)pos);
    }
}

// Presynaptic inhibition (lif.c:38) reads gI of the arriving neuron while the
// step's own deliveries are still landing, so its result depends on the order
// of every earlier arrival.  This variant keeps lif.c's serial order: one
// workgroup walks the arrivals in list order and adds into gE/gI directly.
__global__ void __launch_bounds__(1024) k_deliver_serial(DevBrain d, int32_t head, int32_t rowsUnique) {
    const int32_t A = d.ringCount[head];
    const int32_t *list = d.ring + (int64_t)head * d.ringStride;
    for (int32_t k = 0; k < A; k++) {
        const int32_t pre = list[k];
        float s = d.g.sign[pre] * d.res[pre];
        if (s == 0.f) continue;
        if (d.k.preInh > 0.f && d.drive[pre] > 0.f) {
            const float gi = d.gI[pre];
            const float inh = gi < 0.f ? -gi : 0.f;
            s = s / (1.f + d.k.preInh * inh);
        }
        const uint32_t a = d.g.indptr[pre], e = d.g.indptr[pre + 1];
        __syncthreads();
        if (threadIdx.x == 0) {
            d.res[pre] = d.res[pre] - d.k.depU * d.res[pre];
            atomicAdd(&d.sc->edgeEvents, (unsigned long long)(e - a));
            atomicAdd(&d.sc->synEvents, (unsigned long long)d.g.rowSyn[pre]);
        }
        float *acc = s > 0 ? d.gE : d.gI;
        if (rowsUnique) {
            for (uint32_t j = a + threadIdx.x; j < e; j += blockDim.x) {
                const uint2 ed = d.g.edges[j];
                acc[ed.x] = acc[ed.x] + __uint_as_float(ed.y) * s;
            }
        } else if (threadIdx.x == 0) {
            for (uint32_t j = a; j < e; j++) {
                const uint2 ed = d.g.edges[j];
                acc[ed.x] = acc[ed.x] + __uint_as_float(ed.y) * s;
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ long runs

// The step's (class, target) pairs with more than LONG_RUN arrivals: one
// workgroup each stages the run in LDS, ranks every arrival by its order key,
// and one thread adds the addends in that order onto the conductance, which
// it stores; the pair's count is cleared, so k_membrane sees no arrivals
// there.  Runs longer than LONG_CAP are ranked against global memory.
constexpr int LONG_T = 256;
constexpr int LONG_CAP = 4096;

__global__ void __launch_bounds__(LONG_T) k_long(DevBrain d) {
    __shared__ uint64_t ks[LONG_CAP];
    __shared__ float as[LONG_CAP];
    __shared__ float sorted[LONG_CAP];
    const int tid = threadIdx.x;
    const int32_t nLong = d.sc->nLong;
    for (int32_t e = blockIdx.x; e < nLong; e += gridDim.x) {
        const int32_t ci = d.longList[e];
        const int32_t n = d.cnt[ci];
        const int32_t c = ci >= d.N ? 1 : 0;
        const int32_t i = ci - c * d.N;
        const int64_t r0 = (c ? d.g.E : 0) + d.g.inOff[i];
        const uint64_t *rk = d.runKey + r0;
        const float *ra = d.runAdd + r0;
        // entry 0 is the first arrival, entry q >= 1 is run entry q-1
        const bool inLds = n <= LONG_CAP;
        if (inLds) {
            for (int32_t q = tid; q < n; q += LONG_T) {
                ks[q] = q ? rk[q - 1] : d.key[ci];
                as[q] = q ? ra[q - 1] : d.add[ci];
            }
            __syncthreads();
            for (int32_t q = tid; q < n; q += LONG_T) {
                const uint64_t mine = ks[q];
                int32_t rank = 0;
                for (int32_t m = 0; m < n; m++) rank += ks[m] < mine;
                sorted[rank] = as[q];
            }
            __syncthreads();
            if (tid == 0) {
                float g = c ? d.gI[i] : d.gE[i];
                for (int32_t q = 0; q < n; q++) g = g + sorted[q];
                if (c) d.gI[i] = g; else d.gE[i] = g;
                d.cnt[ci] = 0;
            }
        } else if (tid == 0) {
            // repeated minimum search (keys are distinct)
            float g = c ? d.gI[i] : d.gE[i];
            uint64_t last = 0;
            for (int32_t r = 0; r < n; r++) {
                uint64_t best = ~0ull;
                float ba = 0.f;
                for (int32_t q = 0; q < n; q++) {
                    const uint64_t kk = q ? rk[q - 1] : d.key[ci];
                    if ((r == 0 || kk > last) && kk < best) { best = kk; ba = q ? ra[q - 1] : d.add[ci]; }
                }
                g = g + ba;
                last = best;
            }
            if (c) d.gI[i] = g; else d.gE[i] = g;
            d.cnt[ci] = 0;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ driven

__global__ void __launch_bounds__(1024) k_driven(DevBrain d, int32_t slot) {
    constexpr int T = 1024;
    __shared__ uint32_t leapS[LEAP_LDS_WORDS];
    __shared__ uint32_t ybase[64];
    __shared__ uint32_t draws[LEAP_ROUND];
    __shared__ int32_t scan[64 + T / 64];
    __shared__ uint32_t s_rng;
    __shared__ int32_t s_nbg;
    const int tid = threadIdx.x;
    const int32_t N = d.N;
    leap_load<T>(leapS, d.leap);
    if (tid == 0) {
        // lif.c:47 -- the fractional background accumulator
        float acc = d.sc->bgAcc + d.k.bgEvents;
        const int32_t nbg = (int32_t)acc;
        acc = acc - (float)nbg;
        d.sc->bgAcc = acc;
        s_nbg = nbg;
        s_rng = d.sc->rng;
    }
    __syncthreads();
    uint32_t base = s_rng;
    const int32_t nbg = s_nbg;
    // lif.c:48 -- background draw t is the stream's state at position t+1
    for (int32_t r0 = 0; r0 < nbg; r0 += LEAP_ROUND) {
        const int32_t cnt = min(LEAP_ROUND, nbg - r0);
        leap_round(leapS, ybase, draws, cnt, base);
        for (int32_t t = tid; t < cnt; t += T) atomicAdd(&d.bgCnt[draws[t] % (uint32_t)N], 1);
        base = draws[cnt - 1];
        __syncthreads();
    }

    // lif.c:51-52 -- one draw per non-refractory driven neuron, in list order
    const int32_t nD = d.nDriven;
    const int32_t per = (nD + T - 1) / T;
    const int32_t lo = min(nD, tid * per), hi = min(nD, lo + per);
    int32_t mine = 0;
    for (int32_t q = lo; q < hi; q++)
        if (!(d.refr[d.driven[q]] > 0)) mine++;
    int32_t totalDraws;
    const int32_t off = wg_exclusive_scan<T>(mine, scan, &totalDraws);
    const float forcedRefr = d.k.tRef + d.k.dt;
    int32_t *flag = d.stage;   // stage[] is free until k_membrane; one flag per list slot
    int32_t q = lo, next = off, fired = 0;
    for (int32_t q2 = lo; q2 < hi; q2++) flag[q2] = 0;
    for (int32_t r0 = 0; r0 < totalDraws; r0 += LEAP_ROUND) {
        const int32_t cnt = min(LEAP_ROUND, totalDraws - r0);
        leap_round(leapS, ybase, draws, cnt, base);
        while (next < off + mine && next < r0 + cnt) {
            const int32_t i = d.driven[q];
            if (!(d.refr[i] > 0)) {
                const float pFire = d.drive[i] * d.k.dtS;
                const float u = (float)(draws[next - r0] >> 8) * (1.0f / 16777216.0f);
                if (u < pFire) {
                    d.v[i] = d.k.vReset;
                    d.refr[i] = forcedRefr;
                    d.spikeCount[i] += 1u;
                    d.trace[i] = 1.f;
                    d.adapt[i] = d.adapt[i] + d.k.adaptInc;
                    flag[q] = 1;
                    fired++;
                }
                next++;
            }
            q++;
        }
        base = draws[cnt - 1];
        __syncthreads();
    }
    int32_t totalFired;
    int32_t pos = wg_exclusive_scan<T>(fired, scan, &totalFired);
    int32_t *out = d.ring + (int64_t)slot * d.ringStride;
    for (int32_t q2 = lo; q2 < hi; q2++)
        if (flag[q2]) out[pos++] = d.driven[q2];
    if (tid == 0) {
        d.sc->rng = base;
        d.sc->nDrivenFired = totalFired;
    }
}

// ------------------------------------------------------------------ membrane + threshold

__global__ void __launch_bounds__(256) k_membrane(DevBrain d) {
    __shared__ int32_t wcount[4];
    const int32_t N = d.N;
    const int32_t i = blockIdx.x * 256 + threadIdx.x;
    bool spk = false;
    if (i < N) {
        NeuronState s;
        s.v = d.v[i]; s.gE = d.gE[i]; s.gI = d.gI[i]; s.refr = d.refr[i];
        s.trace = d.trace[i]; s.adapt = d.adapt[i]; s.res = d.res[i];
        const float refr0 = s.refr, trace0 = s.trace, adapt0 = s.adapt, res0 = s.res;
        // this step's arrivals, excitatory then inhibitory accumulator
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const int32_t ci = c * N + i;
            const int32_t n = d.cnt[ci];
            if (n) {
                float g = c ? s.gI : s.gE;
                if (n == 1) {
                    g = g + d.add[ci];
                } else {
                    const int64_t r0 = (c ? d.g.E : 0) + d.g.inOff[i];
                    g = ordered_sum(g, n, d.key[ci], d.add[ci], d.runKey + r0, d.runAdd + r0);
                }
                d.cnt[ci] = 0;
                if (c) s.gI = g; else s.gE = g;
            }
        }
        // background hits land after every delivery (lif.c:48)
        const int32_t nb = d.bgCnt[i];
        if (nb) {
            for (int32_t t = 0; t < nb; t++) s.gE = s.gE + d.k.bgAmp;
            d.bgCnt[i] = 0;
        }
        membrane(d.k, s, d.bias[i]);
        if (crosses(d.k, s, d.thr[i])) {
            spike_effects(d.k, s, d.k.tRef);
            d.spikeCount[i] += 1u;
            spk = true;
        }
        d.v[i] = s.v; d.gE[i] = s.gE; d.gI[i] = s.gI;
        // arrays that often keep their bits (a neuron that is not refractory,
        // never spiked, or without adaptation or depression) are written only
        // when they change
        if (__float_as_uint(s.refr) != __float_as_uint(refr0)) d.refr[i] = s.refr;
        if (__float_as_uint(s.trace) != __float_as_uint(trace0)) d.trace[i] = s.trace;
        if (__float_as_uint(s.adapt) != __float_as_uint(adapt0)) d.adapt[i] = s.adapt;
        if (__float_as_uint(s.res) != __float_as_uint(res0)) d.res[i] = s.res;
    }
    // ascending compaction inside the block: wave ballot, then wave prefix
    const uint64_t mask = __ballot(spk);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) wcount[wave] = __popcll(mask);
    __syncthreads();
    int32_t base = 0;
    for (int w = 0; w < wave; w++) base += wcount[w];
    if (spk) {
        const int32_t pos = base + __popcll(mask & ((1ull << lane) - 1ull));
        d.stage[blockIdx.x * 256 + pos] = i;
    }
    if (threadIdx.x == 0) d.blkCnt[blockIdx.x] = wcount[0] + wcount[1] + wcount[2] + wcount[3];
}

__global__ void __launch_bounds__(256) k_compact(DevBrain d, int32_t slot, int32_t head) {
    __shared__ int32_t red[256];
    const int32_t b = blockIdx.x;
    const int32_t nb = d.blkCnt[b];
    // a block without spikes has nothing to place; only the last one also
    // totals the step.  So the prefix sums cost (spiking blocks + 1) x
    // blocks reads, not blocks x blocks.
    if (nb == 0 && b != (int32_t)gridDim.x - 1) return;
    int32_t sum = 0;
    for (int32_t j = threadIdx.x; j < b; j += 256) sum += d.blkCnt[j];
    red[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    const int32_t before = red[0];
    const int32_t nd = d.sc->nDrivenFired;
    int32_t *out = d.ring + (int64_t)slot * d.ringStride;
    for (int32_t t = threadIdx.x; t < nb; t += 256) out[nd + before + t] = d.stage[b * 256 + t];
    if (b == (int32_t)gridDim.x - 1 && threadIdx.x == 0) {
        const int32_t total = nd + before + nb;
        d.ringCount[slot] = total;
        d.ringCount[head] = 0;               // lif.c:44, after k_deliver consumed it
        d.sc->nLong = 0;
        d.sc->spikes += (unsigned long long)total;
    }
}

// Copies the spike lists of `count` consecutive steps (slots s0, s0+1, ... mod
// nslots) back to back into dst, counts into cnt.
__global__ void __launch_bounds__(256) k_gather(DevBrain d, int32_t s0, int32_t count, int32_t *cnt, int32_t *dst) {
    int64_t off = 0;
    for (int32_t q = 0; q < count; q++) {
        const int32_t slot = (s0 + q) % d.nslots;
        const int32_t n = d.ringCount[slot];
        if (threadIdx.x == 0 && blockIdx.x == 0) cnt[q] = n;
        const int32_t *src = d.ring + (int64_t)slot * d.ringStride;
        for (int32_t t = blockIdx.x * 256 + threadIdx.x; t < n; t += gridDim.x * 256) dst[off + t] = src[t];
        off += n;
    }
}

// The driven draws without background events, up to 4096 driven neurons: each
// of 1024 threads holds at most four consecutive driven-list positions, finds
// its first stream position with a scan of non-refractory flags, and jumps
// there with leapfrog products read from the global tables (M^2048 once past
// 2048, then M^(32 h), then M^(f+1) for the position 2048 a + 32 h + f); its
// later draws are single steps.  Same draws and effects as k_driven.
constexpr int KDF_T = 1024;
constexpr int KDF_DP = 4;

__device__ __forceinline__ void driven_fast(const DevBrain &d, int32_t slot) {
    __shared__ int32_t scan[2][KDF_T / 64];
    const int tid = threadIdx.x;
    const uint32_t rng = d.sc->rng;
    const int32_t nD = d.nDriven, P = (nD + KDF_T - 1) / KDF_T;
    int32_t ii[KDF_DP];
    float pFire[KDF_DP];
    uint32_t nr = 0;   // bit u: driven neuron u of this thread is not refractory and draws
#pragma unroll
    for (int u = 0; u < KDF_DP; u++) {
        const int32_t q = tid * P + u;
        ii[u] = (u < P && q < nD) ? d.driven[q] : -1;
        pFire[u] = 0.f;
        if (ii[u] >= 0) {
            pFire[u] = d.drive[ii[u]] * d.k.dtS;
            if (!(d.refr[ii[u]] > 0)) nr |= 1u << u;
        }
    }
    int32_t n;
    const int32_t off = wg_scan_flags<KDF_T, KDF_DP>(nr, scan[0], &n);
    const float forcedRefr = d.k.tRef + d.k.dt;
    uint32_t fm = 0;   // bit u: neuron u was forced to spike
    if (nr) {
        const uint32_t y = off >= 2048 ? matvec(d.leap + LEAP_FAR * LEAP_WORDS, rng) : rng;
        const int32_t o = off & 2047;
        uint32_t x = matvec(d.leap + (o & 31) * LEAP_WORDS, matvec(d.leap + (32 + (o >> 5)) * LEAP_WORDS, y));
        int32_t j = off;
        bool first = true;
#pragma unroll
        for (int u = 0; u < KDF_DP; u++) {
            if (!((nr >> u) & 1u)) continue;
            if (!first) x = xs32(x);
            first = false;
            const int32_t i = ii[u];
            const float uu = (float)(x >> 8) * (1.0f / 16777216.0f);
            if (uu < pFire[u]) {
                d.v[i] = d.k.vReset;
                d.refr[i] = forcedRefr;
                d.spikeCount[i] += 1u;
                d.trace[i] = 1.f;
                d.adapt[i] = d.adapt[i] + d.k.adaptInc;
                fm |= 1u << u;
            }
            if (j == n - 1) d.sc->rng = x;   // the step's last draw is the stream state after it
            j++;
        }
    }
    int32_t nf;
    int32_t pos = wg_scan_flags<KDF_T, KDF_DP>(fm, scan[1], &nf);
    int32_t *out = d.ring + (int64_t)slot * d.ringStride;
#pragma unroll
    for (int u = 0; u < KDF_DP; u++)
        if ((fm >> u) & 1u) out[pos++] = ii[u];
    if (tid == 0) d.sc->nDrivenFired = nf;
}

// One step's delivery and driven draws in one launch: they share no state
// (the delivery writes the arrival accumulators; the draws read refr and
// drive and write v, refr, trace, adapt, counts and the ring's out slot), so
// the last workgroup draws while the others deliver.
constexpr int DF_BLOCKS = 128;

__global__ void __launch_bounds__(DF_T) k_deliver_driven(DevBrain d, int32_t head, int32_t slot) {
    if ((int32_t)blockIdx.x == DF_BLOCKS) driven_fast(d, slot);
    else deliver_flat(d, head, DF_BLOCKS);
}

// The closed pool loop of the cord ensemble (cordens.bend:227-273), after a
// step's list is final: tally this step's spikes by pool, record the counts,
// and set the drive of the loop's targets for the next step from the pool
// total `delay` steps back, base + gain * total in binary32 (the total of a
// step before the first is 0).
constexpr int LOOP_MAX_POOLS = 8;

__global__ void __launch_bounds__(256) k_pool_loop(DevBrain d, int32_t slot) {
    __shared__ uint32_t cnt[LOOP_MAX_POOLS];
    __shared__ float s_drive;
    const int32_t tid = threadIdx.x;
    if (tid < LOOP_MAX_POOLS) cnt[tid] = 0;
    __syncthreads();
    const int32_t n = d.ringCount[slot];
    const int32_t *list = d.ring + (int64_t)slot * d.ringStride;
    for (int32_t q = tid; q < n; q += 256) {
        const uint32_t p = d.poolOf[list[q]];
        if ((int32_t)p < d.loopPools) atomicAdd(&cnt[p], 1u);
    }
    __syncthreads();
    if (tid == 0) {
        const int32_t k = d.loopState[0] + 1;
        uint32_t total = 0;
        for (int32_t l = 0; l < d.loopPools; l++) total += cnt[l];
        const int32_t r = d.loopState[1];
        if (r < d.loopRecCap) {
            for (int32_t l = 0; l < d.loopPools; l++) d.loopRec[(int64_t)r * d.loopPools + l] = (uint16_t)min(cnt[l], 65535u);
            d.loopState[1] = r + 1;
        }
        d.loopHist[k % d.loopRing] = total;
        const uint32_t back = k - d.loopDelay >= 1 ? d.loopHist[(k - d.loopDelay) % d.loopRing] : 0u;
        const float g = d.loopGain * (float)back;
        s_drive = d.loopBase + g;
        d.loopState[0] = k;
    }
    __syncthreads();
    const float v = s_drive;
    for (int32_t q = tid; q < d.loopNTargets; q += 256) d.drive[d.loopTargets[q]] = v;
}

// ------------------------------------------------------------------ host side

int deliver_grid(const DevBrain &d) {
    (void)d;
    return 1024;
}

fb_status launch_step(const DevBrain &d, int32_t head, int32_t slot, bool serialDeliver, bool rowsUnique,
                      hipStream_t st) {
    if (!serialDeliver && d.k.depU == 0.f && d.k.bgEvents == 0.f && d.nDriven <= KDF_T * KDF_DP) {
        hipLaunchKernelGGL(k_deliver_driven, dim3(DF_BLOCKS + 1), dim3(DF_T), 0, st, d, head, slot);
    } else {
        if (serialDeliver)
            hipLaunchKernelGGL(k_deliver_serial, dim3(1), dim3(1024), 0, st, d, head, (int32_t)rowsUnique);
        else
            hipLaunchKernelGGL(k_deliver, dim3(deliver_grid(d)), dim3(256), 0, st, d, head);
        hipLaunchKernelGGL(k_driven, dim3(1), dim3(1024), 0, st, d, slot);
    }
    if (!serialDeliver) hipLaunchKernelGGL(k_long, dim3(64), dim3(LONG_T), 0, st, d);
    hipLaunchKernelGGL(k_membrane, dim3(d.nblocks), dim3(256), 0, st, d);
    hipLaunchKernelGGL(k_compact, dim3(d.nblocks), dim3(256), 0, st, d, slot, head);
    if (d.poolOf) hipLaunchKernelGGL(k_pool_loop, dim3(1), dim3(256), 0, st, d, slot);
    FB_HIP(hipGetLastError());
    return FB_OK;
}

fb_status launch_gather(const DevBrain &d, int32_t s0, int32_t count, int32_t *cnt, int32_t *dst, hipStream_t st) {
    hipLaunchKernelGGL(k_gather, dim3(64), dim3(256), 0, st, d, s0, count, cnt, dst);
    FB_HIP(hipGetLastError());
    return FB_OK;
}

}  // namespace fb
