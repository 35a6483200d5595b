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
};

struct fb_fv {
    fb_fv_model *m = nullptr;
    float *v[2] = {nullptr, nullptr};
    float *acc = nullptr, *x = nullptr;
    int cur = 0;
    hipStream_t stream = nullptr;
};

namespace fb {

__global__ void __launch_bounds__(256) k_fv_step(int32_t N, const int32_t *inOff, const int32_t *inSrc,
                                                 const float *inW, const float *bias, const float *kdt,
                                                 const float *vIn, float *vOut, float *acc, const float *x) {
    const int32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    float a = 0.f;
    for (int32_t q = inOff[i], e = inOff[i + 1]; q < e; q++) {
        const float r = vIn[inSrc[q]];
        if (r <= 0.f) continue;
        const float w = inW[q] * r;
        a = a + w;
    }
    acc[i] = a;
    const float vi = vIn[i];
    const float d = ((-vi + bias[i]) + a) + x[i];
    vOut[i] = vi + kdt[i] * d;
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

static void fv_model_free(fb_fv_model *m) {
    (void)hipFree(m->inOff); (void)hipFree(m->inSrc); (void)hipFree(m->inW);
    (void)hipFree(m->bias); (void)hipFree(m->kdt);
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
    fb_fv_model *m = new fb_fv_model();
    m->N = N;
    m->E = E;
    m->biasH.assign(bias, bias + N);
    if (hipMalloc(&m->inOff, sizeof(int32_t) * ((size_t)N + 1)) != hipSuccess ||
        hipMalloc(&m->inSrc, sizeof(int32_t) * src.size()) != hipSuccess ||
        hipMalloc(&m->inW, sizeof(float) * w.size()) != hipSuccess ||
        hipMalloc(&m->bias, sizeof(float) * N) != hipSuccess || hipMalloc(&m->kdt, sizeof(float) * N) != hipSuccess) {
        fv_model_free(m);
        set_error("fb_fv_model_create: device allocation failed");
        return FB_ERR_DEVICE;
    }
    FB_HIP(hipMemcpy(m->inOff, off.data(), sizeof(int32_t) * off.size(), hipMemcpyHostToDevice));
    FB_HIP(hipMemcpy(m->inSrc, src.data(), sizeof(int32_t) * src.size(), hipMemcpyHostToDevice));
    FB_HIP(hipMemcpy(m->inW, w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice));
    FB_HIP(hipMemcpy(m->bias, bias, sizeof(float) * N, hipMemcpyHostToDevice));
    FB_HIP(hipMemcpy(m->kdt, kdt.data(), sizeof(float) * N, hipMemcpyHostToDevice));
    *out = m;
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

extern "C" fb_status fb_fv_step(fb_fv *e, int32_t nsteps) {
    FB_REQUIRE(e && nsteps >= 0, FB_ERR_ARG, "fb_fv_step: bad arguments");
    const fb_fv_model *m = e->m;
    const int grid = (m->N + 255) / 256;
    for (int32_t s = 0; s < nsteps; s++) {
        hipLaunchKernelGGL(k_fv_step, dim3(grid), dim3(256), 0, e->stream, m->N, m->inOff, m->inSrc, m->inW, m->bias,
                           m->kdt, e->v[e->cur], e->v[e->cur ^ 1], e->acc, e->x);
        e->cur ^= 1;
    }
    FB_HIP(hipGetLastError());
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
