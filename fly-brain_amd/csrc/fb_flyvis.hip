// fb_flyvis.hip -- the flyvis optic-lobe rate network (Lappalainen et al.
// 2024) that drives the connectome from the eyes: fv_step of the reference's
// kernel (src/wasm/lif.c:82-93) and its host class (src/flyvis.js:3-25).
//
//   acc_i = sum over sources j with v_j > 0 of w_ji * v_j
//   v_i  += kdt_i * (-v_i + bias_i + acc_i + x_i),   kdt = dt / max(tau, dt)
//
// The reference scatters in source order (ascending j, each row in order),
// one binary32 add per edge.  Here each target gathers its in-edges, kept in
// that same (source, row position) order, so every sum is bit-identical; the
// update reads the previous activity from one buffer and writes the other.
// A model (graph, bias, kdt) is shared by its instances (one per eye).
//
// The reverse sweep (FlyVisDiff.backward, src/flyvisdiff.js:128-185) runs
// over a tape that keeps every taped step's pre-step activity and integrator
// argument in HBM (two N-float rows a step; the reference keeps a state every
// four steps and replays, which trades the memory a GPU has for compute).
// One launch per step fuses the fan-out of step t+1 (per source, in row
// order, a binary64 running sum as the reference's), the per-node pass of
// step t and the luminance gradient of step t+1 (per hex column, in map
// order); the two lacc rows alternate between launches.  Arithmetic follows
// JavaScript's: binary64 operations, binary32 stores.
#include <algorithm>
#include <numeric>

#include "fb_internal.h"

struct fb_fv_model {
    int refs = 1;
    int32_t N = 0;
    int64_t E = 0;
    int32_t *inOff = nullptr, *inSrc = nullptr;
    float *inW = nullptr, *bias = nullptr, *kdt = nullptr;
    std::vector<float> biasH;
    // the graph by source, for the reverse sweep's fan-out
    int32_t *indptr = nullptr, *target = nullptr;
    float *weight = nullptr;
    // wave groups: consecutive targets (inGrp) or sources (outGrp) with at
    // most 64 nodes and FV_CH edges, one wave each (a node with more edges
    // is a group of its own)
    int32_t nInGrp = 0, nOutGrp = 0;
    int32_t *inGrp = nullptr, *outGrp = nullptr;
    // the photoreceptor map: x[xNode[q]] = lum[xCol[q]] (the last entry of a
    // node in map order wins), and per hex column its input nodes in map order
    int32_t nIn = 0, nCol = 0, nX = 0;
    int32_t *xNode = nullptr, *xCol = nullptr, *colOff = nullptr, *colNode = nullptr;
};

struct fb_fv {
    fb_fv_model *m = nullptr;
    float *v[2] = {nullptr, nullptr};
    float *acc = nullptr, *x = nullptr, *lum = nullptr;
    int cur = 0;
    hipStream_t stream = nullptr;
    // tape: tapeV[t][N] the activity before taped step t, tapeU[t][N] its
    // integrator argument ((-v + bias) + acc) + x
    float *tapeV = nullptr, *tapeU = nullptr;
    int32_t tapeCap = 0, tapeN = 0;
    // reverse sweep: carried adjoints and device copies of the accumulators
    float *lv = nullptr, *lacc[2] = {nullptr, nullptr}, *gw = nullptr, *gb = nullptr, *gk = nullptr;
    float *dLdV = nullptr, *dLum = nullptr;
    size_t dLdVCap = 0, dLumCap = 0;
};

namespace fb {

// A wave owns 64 consecutive targets, whose in-edges are one contiguous
// range.  It stages that range a chunk at a time: every lane loads and
// multiplies 16 edges at once (independent loads, so the gathers overlap),
// the products go to LDS, and each lane then adds its own target's products
// in order.  A skipped source (v <= 0) stores +0, which leaves a sum that
// starts at +0 unchanged, so the sums are the reference's.
constexpr int FV_CH = 1024;            // edges staged per chunk
constexpr int FV_PER_LANE = FV_CH / 64;

template <bool TAPE>
__global__ void __launch_bounds__(64) k_fv_step(const int32_t *grp, const int32_t *inOff, const int32_t *inSrc,
                                                const float *inW, const float *bias, const float *kdt,
                                                const float *vIn, float *vOut, float *acc, const float *x,
                                                float *tapeV, float *tapeU) {
    __shared__ float prod[FV_CH];
    const int32_t lane = threadIdx.x, i0 = grp[blockIdx.x], iEnd = grp[blockIdx.x + 1], i = i0 + lane;
    const int32_t e0 = inOff[i0], e1 = inOff[iEnd];
    const int32_t my0 = i < iEnd ? inOff[i] : e1, my1 = i < iEnd ? inOff[i + 1] : e1;
    float a = 0.f;
    for (int32_t c = e0; c < e1; c += FV_CH) {
        const int32_t n = min(FV_CH, e1 - c);
        // loads at clamped indices are unconditional, so all of them issue
        // before the first wait; only the stores are guarded
        int32_t src[FV_PER_LANE];
        float w[FV_PER_LANE], r[FV_PER_LANE];
#pragma unroll
        for (int u = 0; u < FV_PER_LANE; u++) {
            const int32_t k = c + min(u * 64 + lane, n - 1);
            src[u] = inSrc[k];
            w[u] = inW[k];
        }
#pragma unroll
        for (int u = 0; u < FV_PER_LANE; u++) r[u] = vIn[src[u]];
#pragma unroll
        for (int u = 0; u < FV_PER_LANE; u++) {
            const int32_t o = u * 64 + lane;
            if (o < n) prod[o] = r[u] <= 0.f ? 0.f : w[u] * r[u];
        }
        __syncthreads();
        int32_t q = max(my0, c);
        const int32_t e = min(my1, c + n);
        for (; q + 8 <= e; q += 8) {        // the reads of eight terms overlap; the adds stay in order
            float t[8];
#pragma unroll
            for (int u = 0; u < 8; u++) t[u] = prod[q - c + u];
#pragma unroll
            for (int u = 0; u < 8; u++) a = a + t[u];
        }
        for (; q < e; q++) a = a + prod[q - c];
        __syncthreads();
    }
    if (i >= iEnd) return;
    acc[i] = a;
    const float vi = vIn[i];
    const float d = ((-vi + bias[i]) + a) + x[i];
    vOut[i] = vi + kdt[i] * d;
    if (TAPE) {
        tapeV[i] = vi;
        tapeU[i] = d;
    }
}

__global__ void __launch_bounds__(256) k_fv_lum(int32_t nX, const int32_t *xNode, const int32_t *xCol,
                                                const float *lum, float *x) {
    const int32_t q = blockIdx.x * 256 + threadIdx.x;
    if (q < nX) x[xNode[q]] = lum[xCol[q]];
}

// One launch of the reverse sweep.  tNext >= 0: the fan-out and luminance
// gradient of step tNext, from laccNext (that step's lacc, complete since the
// previous launch).  tCur >= 0: the per-node pass of step tCur into laccCur.
// A wave owns 64 consecutive sources (one contiguous range of rows): lanes
// take 16 edges of a staged chunk each, update gw of their own edges (a
// word per edge, so in parallel) and put la * w in LDS; each lane then adds
// its own row's terms in row order into a binary64 sum, as the reference's
// loop does (a skipped term stores +0, which leaves a sum started at +0
// unchanged).  Lane j also owns lv[j], gb[j], gk[j] and, in the first
// blocks, hex column j of dLum, so no two lanes write one word.
__global__ void __launch_bounds__(64) k_fv_back(int32_t N, int32_t nCol, int32_t nGrp, const int32_t *grp,
                                                int32_t tNext, int32_t tCur, const int32_t *indptr, const int32_t *target, const float *weight,
                                                const float *kdt, const float *tapeV, const float *tapeU,
                                                const float *dLdV, float *lv, const float *laccNext, float *laccCur,
                                                float *gw, float *gb, float *gk, const int32_t *colOff,
                                                const int32_t *colNode, float *dLum) {
    __shared__ double prod[FV_CH];
    __shared__ float rS[64];
    __shared__ int32_t rowOff[65];
    const int32_t lane = threadIdx.x, col = blockIdx.x * 64 + lane;
    if (tNext >= 0 && dLum != nullptr && col < nCol) {
        float *row = dLum + (size_t)tNext * nCol;
        float a = row[col];
        for (int32_t q = colOff[col], e = colOff[col + 1]; q < e; q++)
            a = (float)((double)a + (double)laccNext[colNode[q]]);
        row[col] = a;
    }
    if ((int32_t)blockIdx.x >= nGrp) return;   // the whole block: no barrier is skipped by part of it
    const int32_t j0 = grp[blockIdx.x], jEnd = grp[blockIdx.x + 1], j = j0 + lane;
    float lvj = j < jEnd ? lv[j] : 0.f;
    if (tNext >= 0) {
        const float r = j < jEnd ? tapeV[(size_t)tNext * N + j] : 0.f;
        rS[lane] = r;
        rowOff[lane] = indptr[min(j, jEnd)];
        if (lane == 0) rowOff[64] = indptr[jEnd];
        __syncthreads();
        const int32_t e0 = rowOff[0], e1 = rowOff[64], my0 = rowOff[lane], my1 = rowOff[lane + 1];
        double sum = 0.0;
        for (int32_t c = e0; c < e1; c += FV_CH) {
            const int32_t n = min(FV_CH, e1 - c);
            // unconditional loads at clamped indices, then the gathers, then
            // the guarded stores (see k_fv_step)
            int32_t tg[FV_PER_LANE];
            float w[FV_PER_LANE], g[FV_PER_LANE], rr[FV_PER_LANE], la[FV_PER_LANE];
#pragma unroll
            for (int u = 0; u < FV_PER_LANE; u++) {
                const int32_t k = c + min(u * 64 + lane, n - 1);
                tg[u] = target[k];
                w[u] = weight[k];
                g[u] = gw[k];
                int32_t lo = 0, hi = 64;        // the source of edge k: rowOff[lo] <= k < rowOff[lo + 1]
                while (hi - lo > 1) {
                    const int32_t mid = (lo + hi) >> 1;
                    if (rowOff[mid] <= k) lo = mid;
                    else hi = mid;
                }
                rr[u] = rS[lo];
            }
#pragma unroll
            for (int u = 0; u < FV_PER_LANE; u++) la[u] = laccNext[tg[u]];
#pragma unroll
            for (int u = 0; u < FV_PER_LANE; u++) {
                const int32_t o = u * 64 + lane;
                if (o < n) {
                    double p = 0.0;
                    if (!(rr[u] <= 0.f) && la[u] != 0.f) {
                        p = (double)la[u] * (double)w[u];
                        gw[c + o] = (float)((double)g[u] + (double)la[u] * (double)rr[u]);
                    }
                    prod[o] = p;
                }
            }
            __syncthreads();
            int32_t q = max(my0, c);
            const int32_t e = min(my1, c + n);
            for (; q + 8 <= e; q += 8) {
                double t[8];
#pragma unroll
                for (int u = 0; u < 8; u++) t[u] = prod[q - c + u];
#pragma unroll
                for (int u = 0; u < 8; u++) sum += t[u];
            }
            for (; q < e; q++) sum += prod[q - c];
            __syncthreads();
        }
        if (j < jEnd && !(r <= 0.f)) lvj = (float)((double)lvj + sum);
    }
    if (j >= jEnd) return;
    if (tCur >= 0) {
        const size_t o = (size_t)tCur * N + j;
        const double l = (double)lvj + (double)dLdV[o];
        if (l == 0.0) {
            laccCur[j] = 0.f;
            lvj = 0.f;
        } else {
            const double k = kdt[j];
            gk[j] = (float)((double)gk[j] + l * (double)tapeU[o]);
            gb[j] = (float)((double)gb[j] + l * k);
            laccCur[j] = (float)(l * k);
            lvj = (float)(l * (1.0 - k));
        }
    }
    lv[j] = lvj;
}

}  // namespace fb

using namespace fb;

#define FB_REQUIRE(cond, code, ...)      \
    do {                                 \
        if (!(cond)) {                   \
            set_error(__VA_ARGS__);      \
            return code;                 \
        }                                \
    } while (0)

// consecutive nodes in groups of at most 64 nodes and FV_CH edges (off:
// per-node edge offsets, N + 1 of them)
static std::vector<int32_t> fv_groups(const std::vector<int32_t> &off, int32_t N) {
    std::vector<int32_t> g{0};
    for (int32_t i = 0; i < N;) {
        int32_t j = i + 1;
        while (j < N && j - i < 64 && off[(size_t)j + 1] - off[(size_t)i] <= FV_CH) j++;
        g.push_back(j);
        i = j;
    }
    return g;
}

static void fv_model_free(fb_fv_model *m) {
    (void)hipFree(m->inOff); (void)hipFree(m->inSrc); (void)hipFree(m->inW);
    (void)hipFree(m->bias); (void)hipFree(m->kdt);
    (void)hipFree(m->indptr); (void)hipFree(m->target); (void)hipFree(m->weight);
    (void)hipFree(m->inGrp); (void)hipFree(m->outGrp);
    (void)hipFree(m->xNode); (void)hipFree(m->xCol); (void)hipFree(m->colOff); (void)hipFree(m->colNode);
    delete m;
}

extern "C" fb_status fb_fv_model_create(int32_t N, int64_t E, const int32_t *indptr, const int32_t *target,
                                        const float *weight, const float *bias, const float *tau, double dt,
                                        fb_fv_model **out) {
    FB_REQUIRE(out && N > 0 && E >= 0 && indptr && bias && tau && dt > 0 && (E == 0 || (target && weight)), FB_ERR_ARG,
               "fb_fv_model_create: bad arguments");
    FB_REQUIRE(indptr[0] == 0 && (int64_t)indptr[N] == E && E < INT32_MAX, FB_ERR_ARG,
               "fb_fv_model_create: indptr must run from 0 to E");
    for (int32_t j = 0; j < N; j++)
        FB_REQUIRE(indptr[j] <= indptr[j + 1], FB_ERR_ARG, "fb_fv_model_create: indptr not monotone at %d", j);
    for (int64_t k = 0; k < E; k++)
        FB_REQUIRE(target[k] >= 0 && target[k] < N, FB_ERR_ARG, "fb_fv_model_create: target out of range at %lld",
                   (long long)k);
    // in-edges per target in (source, row position) order: a stable sort of
    // the source-major edge list by target
    std::vector<int32_t> off((size_t)N + 1, 0), src((size_t)std::max<int64_t>(E, 1)), pos((size_t)N);
    std::vector<float> w((size_t)std::max<int64_t>(E, 1));
    for (int64_t k = 0; k < E; k++) off[(size_t)target[k] + 1]++;
    for (int32_t i = 0; i < N; i++) off[(size_t)i + 1] += off[(size_t)i];
    for (int32_t i = 0; i < N; i++) pos[(size_t)i] = off[(size_t)i];
    for (int32_t j = 0; j < N; j++)
        for (int32_t k = indptr[j]; k < indptr[j + 1]; k++) {
            const int32_t q = pos[(size_t)target[k]]++;
            src[(size_t)q] = j;
            w[(size_t)q] = weight[k];
        }
    // kdt = dt / max(tau, dt) in binary64, stored as binary32 (src/flyvis.js:13)
    std::vector<float> kdt((size_t)N);
    for (int32_t i = 0; i < N; i++) kdt[(size_t)i] = (float)(dt / std::max((double)tau[i], dt));
    const std::vector<int32_t> inGrp = fv_groups(off, N), outGrp = fv_groups(std::vector<int32_t>(indptr, indptr + N + 1), N);
    fb_fv_model *m = new fb_fv_model();
    m->N = N;
    m->E = E;
    m->nInGrp = (int32_t)inGrp.size() - 1;
    m->nOutGrp = (int32_t)outGrp.size() - 1;
    m->biasH.assign(bias, bias + N);
    if (hipMalloc(&m->inOff, sizeof(int32_t) * ((size_t)N + 1)) != hipSuccess ||
        hipMalloc(&m->inSrc, sizeof(int32_t) * src.size()) != hipSuccess ||
        hipMalloc(&m->inW, sizeof(float) * w.size()) != hipSuccess ||
        hipMalloc(&m->bias, sizeof(float) * N) != hipSuccess || hipMalloc(&m->kdt, sizeof(float) * N) != hipSuccess ||
        hipMalloc(&m->indptr, sizeof(int32_t) * ((size_t)N + 1)) != hipSuccess ||
        hipMalloc(&m->target, sizeof(int32_t) * src.size()) != hipSuccess ||
        hipMalloc(&m->weight, sizeof(float) * w.size()) != hipSuccess ||
        hipMalloc(&m->inGrp, sizeof(int32_t) * inGrp.size()) != hipSuccess ||
        hipMalloc(&m->outGrp, sizeof(int32_t) * outGrp.size()) != hipSuccess) {
        fv_model_free(m);
        set_error("fb_fv_model_create: device allocation failed");
        return FB_ERR_DEVICE;
    }
    FB_HIP(hipMemcpy(m->inOff, off.data(), sizeof(int32_t) * off.size(), hipMemcpyHostToDevice));
    FB_HIP(hipMemcpy(m->inSrc, src.data(), sizeof(int32_t) * src.size(), hipMemcpyHostToDevice));
    FB_HIP(hipMemcpy(m->inW, w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice));
    FB_HIP(hipMemcpy(m->bias, bias, sizeof(float) * N, hipMemcpyHostToDevice));
    FB_HIP(hipMemcpy(m->kdt, kdt.data(), sizeof(float) * N, hipMemcpyHostToDevice));
    FB_HIP(hipMemcpy(m->indptr, indptr, sizeof(int32_t) * ((size_t)N + 1), hipMemcpyHostToDevice));
    FB_HIP(hipMemcpy(m->inGrp, inGrp.data(), sizeof(int32_t) * inGrp.size(), hipMemcpyHostToDevice));
    FB_HIP(hipMemcpy(m->outGrp, outGrp.data(), sizeof(int32_t) * outGrp.size(), hipMemcpyHostToDevice));
    if (E > 0) {
        FB_HIP(hipMemcpy(m->target, target, sizeof(int32_t) * (size_t)E, hipMemcpyHostToDevice));
        FB_HIP(hipMemcpy(m->weight, weight, sizeof(float) * (size_t)E, hipMemcpyHostToDevice));
    }
    *out = m;
    return FB_OK;
}

extern "C" fb_status fb_fv_model_set_kdt(fb_fv_model *m, const float *kdt) {
    FB_REQUIRE(m && kdt, FB_ERR_ARG, "fb_fv_model_set_kdt: bad arguments");
    FB_HIP(hipMemcpy(m->kdt, kdt, sizeof(float) * (size_t)m->N, hipMemcpyHostToDevice));
    return FB_OK;
}

extern "C" fb_status fb_fv_model_set_input_map(fb_fv_model *m, int32_t nIn, const int32_t *inNode,
                                               const int32_t *inCol) {
    FB_REQUIRE(m && nIn > 0 && inNode && inCol, FB_ERR_ARG, "fb_fv_model_set_input_map: bad arguments");
    FB_REQUIRE(m->nIn == 0, FB_ERR_ARG, "fb_fv_model_set_input_map: the model already has an input map");
    int32_t nCol = 0;
    for (int32_t q = 0; q < nIn; q++) {
        FB_REQUIRE(inNode[q] >= 0 && inNode[q] < m->N && inCol[q] >= 0, FB_ERR_ARG,
                   "fb_fv_model_set_input_map: entry %d out of range", q);
        nCol = std::max(nCol, inCol[q] + 1);
    }
    // writes: the last entry of each node in map order
    std::vector<int32_t> last((size_t)m->N, -1);
    for (int32_t q = 0; q < nIn; q++) last[(size_t)inNode[q]] = q;
    std::vector<int32_t> xNode, xCol;
    for (int32_t q = 0; q < nIn; q++)
        if (last[(size_t)inNode[q]] == q) {
            xNode.push_back(inNode[q]);
            xCol.push_back(inCol[q]);
        }
    // per column, its entries in map order (a stable counting sort)
    std::vector<int32_t> off((size_t)nCol + 1, 0), node((size_t)nIn);
    for (int32_t q = 0; q < nIn; q++) off[(size_t)inCol[q] + 1]++;
    for (int32_t c = 0; c < nCol; c++) off[(size_t)c + 1] += off[(size_t)c];
    std::vector<int32_t> pos(off.begin(), off.end() - 1);
    for (int32_t q = 0; q < nIn; q++) node[(size_t)pos[(size_t)inCol[q]]++] = inNode[q];
    if (hipMalloc(&m->xNode, sizeof(int32_t) * xNode.size()) != hipSuccess ||
        hipMalloc(&m->xCol, sizeof(int32_t) * xCol.size()) != hipSuccess ||
        hipMalloc(&m->colOff, sizeof(int32_t) * off.size()) != hipSuccess ||
        hipMalloc(&m->colNode, sizeof(int32_t) * node.size()) != hipSuccess) {
        set_error("fb_fv_model_set_input_map: device allocation failed");
        return FB_ERR_DEVICE;
    }
    FB_HIP(hipMemcpy(m->xNode, xNode.data(), sizeof(int32_t) * xNode.size(), hipMemcpyHostToDevice));
    FB_HIP(hipMemcpy(m->xCol, xCol.data(), sizeof(int32_t) * xCol.size(), hipMemcpyHostToDevice));
    FB_HIP(hipMemcpy(m->colOff, off.data(), sizeof(int32_t) * off.size(), hipMemcpyHostToDevice));
    FB_HIP(hipMemcpy(m->colNode, node.data(), sizeof(int32_t) * node.size(), hipMemcpyHostToDevice));
    m->nIn = nIn;
    m->nCol = nCol;
    m->nX = (int32_t)xNode.size();
    return FB_OK;
}

extern "C" fb_status fb_fv_model_release(fb_fv_model *m) {
    FB_REQUIRE(m, FB_ERR_ARG, "fb_fv_model_release: null model");
    if (--m->refs == 0) fv_model_free(m);
    return FB_OK;
}

extern "C" fb_status fb_fv_create(fb_fv_model *m, fb_fv **out) {
    FB_REQUIRE(m && out, FB_ERR_ARG, "fb_fv_create: bad arguments");
    fb_fv *e = new fb_fv();
    e->m = m;
    const size_t N = (size_t)m->N;
    if (hipMalloc(&e->v[0], sizeof(float) * N) != hipSuccess || hipMalloc(&e->v[1], sizeof(float) * N) != hipSuccess ||
        hipMalloc(&e->acc, sizeof(float) * N) != hipSuccess || hipMalloc(&e->x, sizeof(float) * N) != hipSuccess ||
        hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipFree(e->v[0]); (void)hipFree(e->v[1]); (void)hipFree(e->acc); (void)hipFree(e->x);
        delete e;
        set_error("fb_fv_create: device allocation failed");
        return FB_ERR_DEVICE;
    }
    m->refs++;
    // a new instance starts at rest, v = bias (src/flyvis.js:17), with no input
    FB_HIP(hipMemcpy(e->v[0], m->biasH.data(), sizeof(float) * N, hipMemcpyHostToDevice));
    FB_HIP(hipMemset(e->acc, 0, sizeof(float) * N));
    FB_HIP(hipMemset(e->x, 0, sizeof(float) * N));
    *out = e;
    return FB_OK;
}

extern "C" fb_status fb_fv_destroy(fb_fv *e) {
    FB_REQUIRE(e, FB_ERR_ARG, "fb_fv_destroy: null instance");
    (void)hipStreamSynchronize(e->stream);
    (void)hipFree(e->v[0]); (void)hipFree(e->v[1]); (void)hipFree(e->acc); (void)hipFree(e->x);
    (void)hipFree(e->lum); (void)hipFree(e->tapeV); (void)hipFree(e->tapeU); (void)hipFree(e->lv);
    (void)hipFree(e->lacc[0]); (void)hipFree(e->lacc[1]); (void)hipFree(e->gw); (void)hipFree(e->gb);
    (void)hipFree(e->gk); (void)hipFree(e->dLdV); (void)hipFree(e->dLum);
    (void)hipStreamDestroy(e->stream);
    fb_fv_model *m = e->m;
    delete e;
    return fb_fv_model_release(m);
}

extern "C" fb_status fb_fv_set_input(fb_fv *e, const float *x) {
    FB_REQUIRE(e && x, FB_ERR_ARG, "fb_fv_set_input: bad arguments");
    FB_HIP(hipMemcpyAsync(e->x, x, sizeof(float) * (size_t)e->m->N, hipMemcpyHostToDevice, e->stream));
    FB_HIP(hipStreamSynchronize(e->stream));
    return FB_OK;
}

static fb_status fv_run(fb_fv *e, int32_t nsteps, bool tape) {
    const fb_fv_model *m = e->m;
    const int grid = m->nInGrp;
    const size_t N = (size_t)m->N;
    for (int32_t s = 0; s < nsteps; s++) {
        if (tape) {
            hipLaunchKernelGGL(k_fv_step<true>, dim3(grid), dim3(64), 0, e->stream, m->inGrp, m->inOff, m->inSrc, m->inW,
                               m->bias, m->kdt, e->v[e->cur], e->v[e->cur ^ 1], e->acc, e->x,
                               e->tapeV + (size_t)e->tapeN * N, e->tapeU + (size_t)e->tapeN * N);
            e->tapeN++;
        } else {
            hipLaunchKernelGGL(k_fv_step<false>, dim3(grid), dim3(64), 0, e->stream, m->inGrp, m->inOff, m->inSrc,
                               m->inW, m->bias, m->kdt, e->v[e->cur], e->v[e->cur ^ 1], e->acc, e->x, nullptr,
                               nullptr);
        }
        e->cur ^= 1;
    }
    FB_HIP(hipGetLastError());
    return FB_OK;
}

extern "C" fb_status fb_fv_step(fb_fv *e, int32_t nsteps) {
    FB_REQUIRE(e && nsteps >= 0, FB_ERR_ARG, "fb_fv_step: bad arguments");
    return fv_run(e, nsteps, false);
}

extern "C" fb_status fb_fv_set_lum(fb_fv *e, const float *lum) {
    FB_REQUIRE(e && lum, FB_ERR_ARG, "fb_fv_set_lum: bad arguments");
    const fb_fv_model *m = e->m;
    FB_REQUIRE(m->nIn > 0, FB_ERR_ARG, "fb_fv_set_lum: the model has no input map");
    if (!e->lum && hipMalloc(&e->lum, sizeof(float) * (size_t)m->nCol) != hipSuccess) {
        set_error("fb_fv_set_lum: device allocation failed");
        return FB_ERR_DEVICE;
    }
    FB_HIP(hipMemcpyAsync(e->lum, lum, sizeof(float) * (size_t)m->nCol, hipMemcpyHostToDevice, e->stream));
    hipLaunchKernelGGL(k_fv_lum, dim3((m->nX + 255) / 256), dim3(256), 0, e->stream, m->nX, m->xNode, m->xCol, e->lum,
                       e->x);
    FB_HIP(hipGetLastError());
    FB_HIP(hipStreamSynchronize(e->stream));
    return FB_OK;
}

extern "C" fb_status fb_fv_tape_start(fb_fv *e, int32_t steps) {
    FB_REQUIRE(e && steps > 0, FB_ERR_ARG, "fb_fv_tape_start: bad arguments");
    if (steps > e->tapeCap) {
        FB_HIP(hipStreamSynchronize(e->stream));
        (void)hipFree(e->tapeV);
        (void)hipFree(e->tapeU);
        e->tapeV = e->tapeU = nullptr;
        e->tapeCap = 0;
        const size_t bytes = sizeof(float) * (size_t)steps * (size_t)e->m->N;
        if (hipMalloc(&e->tapeV, bytes) != hipSuccess || hipMalloc(&e->tapeU, bytes) != hipSuccess) {
            (void)hipFree(e->tapeV);
            e->tapeV = nullptr;
            set_error("fb_fv_tape_start: device allocation of %d steps failed", steps);
            return FB_ERR_DEVICE;
        }
        e->tapeCap = steps;
    }
    e->tapeN = 0;
    return FB_OK;
}

extern "C" fb_status fb_fv_step_taped(fb_fv *e, int32_t nsteps) {
    FB_REQUIRE(e && nsteps >= 0, FB_ERR_ARG, "fb_fv_step_taped: bad arguments");
    FB_REQUIRE(e->tapeCap > 0, FB_ERR_ARG, "fb_fv_step_taped: no tape (call fb_fv_tape_start)");
    FB_REQUIRE((int64_t)e->tapeN + nsteps <= e->tapeCap, FB_ERR_ARG,
               "fb_fv_step_taped: the tape holds %d steps, %d are taped", e->tapeCap, e->tapeN);
    return fv_run(e, nsteps, true);
}

extern "C" fb_status fb_fv_tape_len(fb_fv *e, int32_t *n) {
    FB_REQUIRE(e && n, FB_ERR_ARG, "fb_fv_tape_len: bad arguments");
    *n = e->tapeN;
    return FB_OK;
}

extern "C" fb_status fb_fv_model_size(const fb_fv_model *m, int32_t *N, int64_t *E) {
    FB_REQUIRE(m && N && E, FB_ERR_ARG, "fb_fv_model_size: bad arguments");
    *N = m->N;
    *E = m->E;
    return FB_OK;
}

extern "C" fb_status fb_fv_size(const fb_fv *e, int32_t *N, int64_t *E, int32_t *nCol) {
    FB_REQUIRE(e && N && E && nCol, FB_ERR_ARG, "fb_fv_size: bad arguments");
    *N = e->m->N;
    *E = e->m->E;
    *nCol = e->m->nCol;
    return FB_OK;
}

static bool fv_grow(float **p, size_t *cap, size_t n) {
    if (n <= *cap) return true;
    (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    if (hipMalloc(p, sizeof(float) * n) != hipSuccess) return false;
    *cap = n;
    return true;
}

extern "C" fb_status fb_fv_backward(fb_fv *e, const float *dLdV, float *gW, float *gBias, float *gKdt,
                                    float *dLdLum) {
    FB_REQUIRE(e && dLdV && gW && gBias && gKdt, FB_ERR_ARG, "fb_fv_backward: bad arguments");
    const fb_fv_model *m = e->m;
    FB_REQUIRE(!dLdLum || m->nIn > 0, FB_ERR_ARG, "fb_fv_backward: a luminance gradient needs the input map");
    const int32_t n = e->tapeN;
    if (n == 0) return FB_OK;
    const size_t N = (size_t)m->N, E = (size_t)m->E, nCol = (size_t)m->nCol;
    FB_HIP(hipStreamSynchronize(e->stream));
    if (!e->lv) {
        if (hipMalloc(&e->lv, sizeof(float) * N) != hipSuccess || hipMalloc(&e->lacc[0], sizeof(float) * N) != hipSuccess ||
            hipMalloc(&e->lacc[1], sizeof(float) * N) != hipSuccess ||
            hipMalloc(&e->gw, sizeof(float) * std::max<size_t>(E, 1)) != hipSuccess ||
            hipMalloc(&e->gb, sizeof(float) * N) != hipSuccess || hipMalloc(&e->gk, sizeof(float) * N) != hipSuccess) {
            set_error("fb_fv_backward: device allocation failed");
            return FB_ERR_DEVICE;
        }
    }
    if (!fv_grow(&e->dLdV, &e->dLdVCap, (size_t)n * N) || (dLdLum && !fv_grow(&e->dLum, &e->dLumCap, (size_t)n * nCol))) {
        set_error("fb_fv_backward: device allocation failed");
        return FB_ERR_DEVICE;
    }
    hipStream_t st = e->stream;
    FB_HIP(hipMemcpyAsync(e->dLdV, dLdV, sizeof(float) * (size_t)n * N, hipMemcpyHostToDevice, st));
    if (E > 0) FB_HIP(hipMemcpyAsync(e->gw, gW, sizeof(float) * E, hipMemcpyHostToDevice, st));
    FB_HIP(hipMemcpyAsync(e->gb, gBias, sizeof(float) * N, hipMemcpyHostToDevice, st));
    FB_HIP(hipMemcpyAsync(e->gk, gKdt, sizeof(float) * N, hipMemcpyHostToDevice, st));
    if (dLdLum) FB_HIP(hipMemcpyAsync(e->dLum, dLdLum, sizeof(float) * (size_t)n * nCol, hipMemcpyHostToDevice, st));
    FB_HIP(hipMemsetAsync(e->lv, 0, sizeof(float) * N, st));
    float *dLum = dLdLum ? e->dLum : nullptr;
    const int grid = std::max(m->nOutGrp, dLdLum ? (int)((nCol + 63) / 64) : 0);
    // launch k handles the fan-out of step n-k and the per-node pass of step
    // n-1-k; step t's lacc lives in lacc[t & 1]
    for (int32_t t = n - 1; t >= -1; t--) {
        const int32_t tNext = t + 1 < n ? t + 1 : -1;
        const float *laccNext = tNext >= 0 ? e->lacc[tNext & 1] : nullptr;
        float *laccCur = t >= 0 ? e->lacc[t & 1] : nullptr;
        hipLaunchKernelGGL(k_fv_back, dim3(grid), dim3(64), 0, st, m->N, (int32_t)nCol, m->nOutGrp, m->outGrp, tNext, t,
                           m->indptr, m->target,
                           m->weight, m->kdt, e->tapeV, e->tapeU, e->dLdV, e->lv, laccNext, laccCur, e->gw, e->gb, e->gk,
                           m->colOff, m->colNode, dLum);
    }
    FB_HIP(hipGetLastError());
    if (E > 0) FB_HIP(hipMemcpyAsync(gW, e->gw, sizeof(float) * E, hipMemcpyDeviceToHost, st));
    FB_HIP(hipMemcpyAsync(gBias, e->gb, sizeof(float) * N, hipMemcpyDeviceToHost, st));
    FB_HIP(hipMemcpyAsync(gKdt, e->gk, sizeof(float) * N, hipMemcpyDeviceToHost, st));
    if (dLdLum) FB_HIP(hipMemcpyAsync(dLdLum, e->dLum, sizeof(float) * (size_t)n * nCol, hipMemcpyDeviceToHost, st));
    FB_HIP(hipStreamSynchronize(st));
    return FB_OK;
}

extern "C" fb_status fb_fv_get(fb_fv *e, int32_t which, float *dst) {
    FB_REQUIRE(e && dst && (which == FB_FV_V || which == FB_FV_ACC), FB_ERR_ARG, "fb_fv_get: bad arguments");
    FB_HIP(hipMemcpyAsync(dst, which == FB_FV_V ? e->v[e->cur] : e->acc, sizeof(float) * (size_t)e->m->N,
                          hipMemcpyDeviceToHost, e->stream));
    FB_HIP(hipStreamSynchronize(e->stream));
    return FB_OK;
}

extern "C" fb_status fb_fv_set_v(fb_fv *e, const float *v) {
    FB_REQUIRE(e && v, FB_ERR_ARG, "fb_fv_set_v: bad arguments");
    FB_HIP(hipMemcpyAsync(e->v[e->cur], v, sizeof(float) * (size_t)e->m->N, hipMemcpyHostToDevice, e->stream));
    FB_HIP(hipStreamSynchronize(e->stream));
    return FB_OK;
}

extern "C" fb_status fb_fv_reset(fb_fv *e) {
    FB_REQUIRE(e, FB_ERR_ARG, "fb_fv_reset: null instance");
    return fb_fv_set_v(e, e->m->biasH.data());
}
