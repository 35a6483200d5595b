// fb_napi.cc -- N-API binding of libflybrain for Node (the reference's host
// runtime).  Functions map one-to-one onto include/flybrain.h; handles are
// napi externals, arrays are TypedArrays the caller owns (copied in or out,
// never retained).  Errors become JS exceptions carrying fb_last_error().
// lifhip.js builds the LIFWasm-compatible class on top of these functions,
// flyviship.js the FlyVis / FlyVisDiff classes on the fv* ones.
#include <node_api.h>

#include <cstdint>
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "flybrain.h"

namespace {

#define NAPI_OK(call)                                                          \
    do {                                                                       \
        if ((call) != napi_ok) {                                               \
            napi_throw_error(env, nullptr, "flybrain napi: internal N-API call failed"); \
            return nullptr;                                                    \
        }                                                                      \
    } while (0)

napi_value throw_fb(napi_env env, fb_status st, const char *what) {
    std::string m = std::string(what) + ": flybrain error " + std::to_string((int)st) + ": " + fb_last_error();
    napi_throw_error(env, nullptr, m.c_str());
    return nullptr;
}

#define FB_OK_OR_THROW(call, what)                 \
    do {                                           \
        fb_status _s = (call);                     \
        if (_s != FB_OK) return throw_fb(env, _s, what); \
    } while (0)

struct Args {
    napi_value v[8];
    size_t n = 8;
};

bool get_args(napi_env env, napi_callback_info info, Args &a, size_t need) {
    if (napi_get_cb_info(env, info, &a.n, a.v, nullptr, nullptr) != napi_ok) return false;
    if (a.n < need) {
        napi_throw_type_error(env, nullptr, "flybrain napi: too few arguments");
        return false;
    }
    return true;
}

void *ext(napi_env env, napi_value v) {
    void *p = nullptr;
    if (napi_get_value_external(env, v, &p) != napi_ok) {
        napi_throw_type_error(env, nullptr, "flybrain napi: expected a handle");
        return nullptr;
    }
    return p;
}

double num(napi_env env, napi_value v) {
    double d = 0;
    napi_get_value_double(env, v, &d);
    return d;
}

int32_t i32(napi_env env, napi_value v) {
    int32_t d = 0;
    napi_get_value_int32(env, v, &d);
    return d;
}

// data pointer and element count of a TypedArray (null / undefined -> null)
template <class T>
bool typed(napi_env env, napi_value v, T **data, size_t *len, bool optional = false) {
    napi_valuetype t;
    napi_typeof(env, v, &t);
    if (optional && (t == napi_null || t == napi_undefined)) {
        *data = nullptr;
        *len = 0;
        return true;
    }
    bool is = false;
    napi_is_typedarray(env, v, &is);
    if (!is) {
        napi_throw_type_error(env, nullptr, "flybrain napi: expected a TypedArray");
        return false;
    }
    napi_typedarray_type tt;
    size_t n = 0;
    void *p = nullptr;
    napi_value ab;
    size_t off = 0;
    napi_get_typedarray_info(env, v, &tt, &n, &p, &ab, &off);
    size_t esz = 1;
    switch (tt) {
        case napi_int8_array: case napi_uint8_array: case napi_uint8_clamped_array: esz = 1; break;
        case napi_int16_array: case napi_uint16_array: esz = 2; break;
        case napi_int32_array: case napi_uint32_array: case napi_float32_array: esz = 4; break;
        default: esz = 8; break;
    }
    if (esz != sizeof(T)) {
        napi_throw_type_error(env, nullptr, "flybrain napi: TypedArray element size mismatch");
        return false;
    }
    *data = (T *)p;
    *len = n;
    return true;
}

// a caller's TypedArray must hold at least `want` elements: the C side copies
// exactly that many into or out of it
bool long_enough(napi_env env, size_t have, size_t want, const char *what) {
    if (have >= want) return true;
    std::string m = std::string(what) + ": array of " + std::to_string(have) + " elements, " +
                    std::to_string(want) + " needed";
    napi_throw_range_error(env, nullptr, m.c_str());
    return false;
}

bool brain_n(napi_env env, fb_brain *b, size_t *n) {
    int32_t N = 0;
    fb_status st = fb_brain_size(b, &N);
    if (st != FB_OK) { throw_fb(env, st, "size"); return false; }
    *n = (size_t)N;
    return true;
}

bool fv_dims(napi_env env, fb_fv *e, size_t *N, size_t *E, size_t *nCol, size_t *tape) {
    int32_t n = 0, c = 0, t = 0;
    int64_t ne = 0;
    fb_status st = fb_fv_size(e, &n, &ne, &c);
    if (st == FB_OK) st = fb_fv_tape_len(e, &t);
    if (st != FB_OK) { throw_fb(env, st, "fvSize"); return false; }
    *N = (size_t)n; *E = (size_t)ne; *nCol = (size_t)c; *tape = (size_t)t;
    return true;
}

napi_value make_i32(napi_env env, int32_t x) {
    napi_value r;
    napi_create_int32(env, x, &r);
    return r;
}

napi_value make_num(napi_env env, double x) {
    napi_value r;
    napi_create_double(env, x, &r);
    return r;
}

napi_value make_ext(napi_env env, void *p) {
    napi_value r;
    napi_create_external(env, p, nullptr, nullptr, &r);
    return r;
}

// ---------------------------------------------------------------- graph / brain

napi_value graph_create(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 6)) return nullptr;
    const int32_t N = i32(env, a.v[0]);
    uint32_t *ip, *ix;
    float *w, *sg;
    int32_t *cn;
    size_t nip, nix, nw, ns, nc;
    if (!typed(env, a.v[1], &ip, &nip) || !typed(env, a.v[2], &ix, &nix) || !typed(env, a.v[3], &w, &nw) ||
        !typed(env, a.v[4], &sg, &ns) || !typed(env, a.v[5], &cn, &nc, true))
        return nullptr;
    if (nip != (size_t)N + 1 || nw != nix || ns != (size_t)N || (cn && nc != nix)) {
        napi_throw_range_error(env, nullptr, "graphCreate: array sizes do not match N and E");
        return nullptr;
    }
    fb_graph *g = nullptr;
    FB_OK_OR_THROW(fb_graph_create(N, (int64_t)nix, ip, ix, w, sg, cn, &g), "graphCreate");
    return make_ext(env, g);
}

napi_value graph_release(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 1)) return nullptr;
    fb_graph *g = (fb_graph *)ext(env, a.v[0]);
    if (!g) return nullptr;
    FB_OK_OR_THROW(fb_graph_release(g), "graphRelease");
    return nullptr;
}

// params: Float64Array in fb_params field order
bool read_params(napi_env env, napi_value v, fb_params *p) {
    double *d;
    size_t n;
    if (!typed(env, v, &d, &n)) return false;
    if (n * sizeof(double) != sizeof(fb_params)) {
        napi_throw_range_error(env, nullptr, "params: wrong length for fb_params");
        return false;
    }
    std::memcpy(p, d, sizeof(fb_params));
    return true;
}

napi_value brain_create(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 3)) return nullptr;
    fb_graph *g = (fb_graph *)ext(env, a.v[0]);
    if (!g) return nullptr;
    fb_params p;
    if (!read_params(env, a.v[1], &p)) return nullptr;
    fb_brain *b = nullptr;
    FB_OK_OR_THROW(fb_brain_create(g, &p, num(env, a.v[2]), &b), "brainCreate");
    return make_ext(env, b);
}

#define BRAIN_OR_NULL(idx)                        \
    fb_brain *b = (fb_brain *)ext(env, a.v[idx]); \
    if (!b) return nullptr

napi_value brain_destroy(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 1)) return nullptr;
    BRAIN_OR_NULL(0);
    FB_OK_OR_THROW(fb_brain_destroy(b), "destroy");
    return nullptr;
}

napi_value reset(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 1)) return nullptr;
    BRAIN_OR_NULL(0);
    FB_OK_OR_THROW(fb_reset(b), "reset");
    return nullptr;
}

napi_value set_params(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 2)) return nullptr;
    BRAIN_OR_NULL(0);
    fb_params p;
    if (!read_params(env, a.v[1], &p)) return nullptr;
    FB_OK_OR_THROW(fb_set_params(b, &p), "setParams");
    return nullptr;
}

// setDrive(b, Int32Array idx, hz) / setBias / pulse share the shape
template <fb_status (*F)(fb_brain *, const int32_t *, int32_t, double)>
napi_value list_setter(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 3)) return nullptr;
    BRAIN_OR_NULL(0);
    int32_t *ix;
    size_t n;
    if (!typed(env, a.v[1], &ix, &n)) return nullptr;
    FB_OK_OR_THROW(F(b, ix, (int32_t)n, num(env, a.v[2])), "setter");
    return nullptr;
}

napi_value set_drive_one(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 3)) return nullptr;
    BRAIN_OR_NULL(0);
    FB_OK_OR_THROW(fb_set_drive_one(b, i32(env, a.v[1]), num(env, a.v[2])), "setDriveOne");
    return nullptr;
}

napi_value set_thr(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 3)) return nullptr;
    BRAIN_OR_NULL(0);
    FB_OK_OR_THROW(fb_set_thr(b, i32(env, a.v[1]), num(env, a.v[2])), "setThr");
    return nullptr;
}

napi_value add_g(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 4)) return nullptr;
    BRAIN_OR_NULL(0);
    FB_OK_OR_THROW(fb_add_g(b, i32(env, a.v[1]), num(env, a.v[2]), num(env, a.v[3])), "addG");
    return nullptr;
}

// applyDeltas(b, Int32Array idx, Int32Array kind, Float64Array val, n): the first n queued setter calls
napi_value apply_deltas(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 5)) return nullptr;
    BRAIN_OR_NULL(0);
    int32_t *ix, *kd;
    double *vl;
    size_t ni, nk, nv;
    if (!typed(env, a.v[1], &ix, &ni) || !typed(env, a.v[2], &kd, &nk) || !typed(env, a.v[3], &vl, &nv)) return nullptr;
    const int32_t n = i32(env, a.v[4]);
    if (n < 0 || (size_t)n > ni || (size_t)n > nk || (size_t)n > nv) {
        napi_throw_range_error(env, nullptr, "flybrain napi: applyDeltas count exceeds the arrays");
        return nullptr;
    }
    FB_OK_OR_THROW(fb_apply_deltas(b, n, ix, kd, vl), "applyDeltas");
    return nullptr;
}

napi_value set_background(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 3)) return nullptr;
    BRAIN_OR_NULL(0);
    FB_OK_OR_THROW(fb_set_background(b, num(env, a.v[1]), num(env, a.v[2])), "setBackground");
    return nullptr;
}

// step(b, nsteps, engine) -> spikes of the last step
napi_value step(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 3)) return nullptr;
    BRAIN_OR_NULL(0);
    int32_t nf = 0;
    FB_OK_OR_THROW(fb_step(b, i32(env, a.v[1]), i32(env, a.v[2]), &nf), "step");
    return make_i32(env, nf);
}

// fired(b, Int32Array out) -> count (lif.c order; at most out.length copied)
napi_value fired(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 2)) return nullptr;
    BRAIN_OR_NULL(0);
    int32_t *out;
    size_t cap;
    if (!typed(env, a.v[1], &out, &cap)) return nullptr;
    int32_t n = 0;
    FB_OK_OR_THROW(fb_fired(b, out, (int32_t)cap, &n), "fired");
    return make_i32(env, n);
}

// getArray(b, which, TypedArray out) / setArray(b, which, TypedArray in): N x 4 bytes
napi_value get_array(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 3)) return nullptr;
    BRAIN_OR_NULL(0);
    float *out;
    size_t n, N;
    if (!typed(env, a.v[2], &out, &n) || !brain_n(env, b, &N) || !long_enough(env, n, N, "getArray")) return nullptr;
    FB_OK_OR_THROW(fb_get_array(b, i32(env, a.v[1]), out), "getArray");
    return nullptr;
}

napi_value set_array(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 3)) return nullptr;
    BRAIN_OR_NULL(0);
    float *in;
    size_t n, N;
    if (!typed(env, a.v[2], &in, &n) || !brain_n(env, b, &N) || !long_enough(env, n, N, "setArray")) return nullptr;
    FB_OK_OR_THROW(fb_set_array(b, i32(env, a.v[1]), in), "setArray");
    return nullptr;
}

// scalars(b) -> [rng, head, bgAcc, tMs]
napi_value scalars(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 1)) return nullptr;
    BRAIN_OR_NULL(0);
    uint32_t rng = 0;
    int32_t head = 0;
    float bg = 0;
    double t = 0;
    FB_OK_OR_THROW(fb_brain_scalars(b, &rng, &head, &bg, &t), "scalars");
    napi_value arr;
    NAPI_OK(napi_create_array_with_length(env, 4, &arr));
    napi_set_element(env, arr, 0, make_num(env, rng));
    napi_set_element(env, arr, 1, make_i32(env, head));
    napi_set_element(env, arr, 2, make_num(env, bg));
    napi_set_element(env, arr, 3, make_num(env, t));
    return arr;
}

// counters(b) -> [edgeEvents, synapseEvents, spikes, steps]
napi_value counters(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 1)) return nullptr;
    BRAIN_OR_NULL(0);
    int64_t c[4] = {0, 0, 0, 0};
    FB_OK_OR_THROW(fb_brain_counters(b, &c[0], &c[1], &c[2], &c[3]), "counters");
    napi_value arr;
    NAPI_OK(napi_create_array_with_length(env, 4, &arr));
    for (int k = 0; k < 4; k++) napi_set_element(env, arr, k, make_num(env, (double)c[k]));
    return arr;
}

// ---------------------------------------------------------------- async runs

struct RunWork {
    fb_brain *b;
    int32_t nsteps, engine, nf = 0;
    fb_status st = FB_OK;
    std::string err;
    napi_deferred deferred;
    napi_async_work work;
};

void run_execute(napi_env, void *data) {
    RunWork *w = (RunWork *)data;
    w->st = fb_step(w->b, w->nsteps, w->engine, &w->nf);
    if (w->st == FB_OK) w->st = fb_synchronize(w->b);
    if (w->st != FB_OK) w->err = fb_last_error();
}

void run_complete(napi_env env, napi_status, void *data) {
    RunWork *w = (RunWork *)data;
    if (w->st == FB_OK) {
        napi_resolve_deferred(env, w->deferred, make_i32(env, w->nf));
    } else {
        napi_value msg, e;
        std::string m = "run: flybrain error " + std::to_string((int)w->st) + ": " + w->err;
        napi_create_string_utf8(env, m.c_str(), m.size(), &msg);
        napi_create_error(env, nullptr, msg, &e);
        napi_reject_deferred(env, w->deferred, e);
    }
    napi_delete_async_work(env, w->work);
    delete w;
}

// runAsync(b, nsteps, engine) -> Promise<spikes of the last step>; the steps
// run off the JS thread
napi_value run_async(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 3)) return nullptr;
    BRAIN_OR_NULL(0);
    RunWork *w = new RunWork();
    w->b = b;
    w->nsteps = i32(env, a.v[1]);
    w->engine = i32(env, a.v[2]);
    napi_value promise, name;
    NAPI_OK(napi_create_promise(env, &w->deferred, &promise));
    NAPI_OK(napi_create_string_utf8(env, "flybrain.run", NAPI_AUTO_LENGTH, &name));
    NAPI_OK(napi_create_async_work(env, nullptr, name, run_execute, run_complete, w, &w->work));
    NAPI_OK(napi_queue_async_work(env, w->work));
    return promise;
}

napi_value device_info(napi_env env, napi_callback_info) {
    char name[256] = {0};
    int32_t ncu = 0;
    FB_OK_OR_THROW(fb_device_info(name, (int32_t)sizeof name, &ncu), "deviceInfo");
    napi_value obj, s;
    NAPI_OK(napi_create_object(env, &obj));
    NAPI_OK(napi_create_string_utf8(env, name, NAPI_AUTO_LENGTH, &s));
    napi_set_named_property(env, obj, "name", s);
    napi_set_named_property(env, obj, "computeUnits", make_i32(env, ncu));
    return obj;
}

napi_value params_length(napi_env env, napi_callback_info) {
    return make_i32(env, (int32_t)(sizeof(fb_params) / sizeof(double)));
}

// ---------------------------------------------------------------- optic lobe (fb_fv_*)

napi_value fv_model_create(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 7)) return nullptr;
    const int32_t N = i32(env, a.v[0]);
    int32_t *ip, *tg;
    float *w, *bias, *tau;
    size_t nip, ntg, nw, nb, nt;
    if (!typed(env, a.v[1], &ip, &nip) || !typed(env, a.v[2], &tg, &ntg) || !typed(env, a.v[3], &w, &nw) ||
        !typed(env, a.v[4], &bias, &nb) || !typed(env, a.v[5], &tau, &nt))
        return nullptr;
    if (nip != (size_t)N + 1 || nw != ntg || nb != (size_t)N || nt != (size_t)N) {
        napi_throw_range_error(env, nullptr, "fvModelCreate: array sizes do not match N and E");
        return nullptr;
    }
    fb_fv_model *m = nullptr;
    FB_OK_OR_THROW(fb_fv_model_create(N, (int64_t)ntg, ip, tg, w, bias, tau, num(env, a.v[6]), &m), "fvModelCreate");
    return make_ext(env, m);
}

napi_value fv_model_set_kdt(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 2)) return nullptr;
    fb_fv_model *m = (fb_fv_model *)ext(env, a.v[0]);
    float *k;
    size_t n;
    if (!m || !typed(env, a.v[1], &k, &n)) return nullptr;
    int32_t N = 0;
    int64_t E = 0;
    FB_OK_OR_THROW(fb_fv_model_size(m, &N, &E), "fvModelSetKdt");
    if (!long_enough(env, n, (size_t)N, "fvModelSetKdt")) return nullptr;
    FB_OK_OR_THROW(fb_fv_model_set_kdt(m, k), "fvModelSetKdt");
    return nullptr;
}

napi_value fv_model_set_input_map(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 3)) return nullptr;
    fb_fv_model *m = (fb_fv_model *)ext(env, a.v[0]);
    int32_t *nd, *cl;
    size_t nn, nc;
    if (!m || !typed(env, a.v[1], &nd, &nn) || !typed(env, a.v[2], &cl, &nc)) return nullptr;
    if (nn != nc) {
        napi_throw_range_error(env, nullptr, "fvModelSetInputMap: node and column lists differ in length");
        return nullptr;
    }
    FB_OK_OR_THROW(fb_fv_model_set_input_map(m, (int32_t)nn, nd, cl), "fvModelSetInputMap");
    return nullptr;
}

napi_value fv_model_release(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 1)) return nullptr;
    fb_fv_model *m = (fb_fv_model *)ext(env, a.v[0]);
    if (!m) return nullptr;
    FB_OK_OR_THROW(fb_fv_model_release(m), "fvModelRelease");
    return nullptr;
}

napi_value fv_create(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 1)) return nullptr;
    fb_fv_model *m = (fb_fv_model *)ext(env, a.v[0]);
    if (!m) return nullptr;
    fb_fv *e = nullptr;
    FB_OK_OR_THROW(fb_fv_create(m, &e), "fvCreate");
    return make_ext(env, e);
}

// fv calls of the form f(e), f(e, int), f(e, array)
template <fb_status (*F)(fb_fv *)>
napi_value fv_call0(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 1)) return nullptr;
    fb_fv *e = (fb_fv *)ext(env, a.v[0]);
    if (!e) return nullptr;
    FB_OK_OR_THROW(F(e), "fv");
    return nullptr;
}

template <fb_status (*F)(fb_fv *, int32_t)>
napi_value fv_call_int(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 2)) return nullptr;
    fb_fv *e = (fb_fv *)ext(env, a.v[0]);
    if (!e) return nullptr;
    FB_OK_OR_THROW(F(e, i32(env, a.v[1])), "fv");
    return nullptr;
}

// PerCol: the array is one value per luminance column, else one per node
template <fb_status (*F)(fb_fv *, const float *), bool PerCol>
napi_value fv_call_arr(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 2)) return nullptr;
    fb_fv *e = (fb_fv *)ext(env, a.v[0]);
    float *x;
    size_t n, N, E, nCol, tape;
    if (!e || !typed(env, a.v[1], &x, &n) || !fv_dims(env, e, &N, &E, &nCol, &tape) ||
        !long_enough(env, n, PerCol ? nCol : N, PerCol ? "fvSetLum" : "fvSetInput/fvSetV"))
        return nullptr;
    FB_OK_OR_THROW(F(e, x), "fv");
    return nullptr;
}

napi_value fv_get(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 3)) return nullptr;
    fb_fv *e = (fb_fv *)ext(env, a.v[0]);
    float *dst;
    size_t n, N, E, nCol, tape;
    if (!e || !typed(env, a.v[2], &dst, &n) || !fv_dims(env, e, &N, &E, &nCol, &tape) ||
        !long_enough(env, n, N, "fvGet"))
        return nullptr;
    FB_OK_OR_THROW(fb_fv_get(e, i32(env, a.v[1]), dst), "fvGet");
    return nullptr;
}

napi_value fv_tape_len(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 1)) return nullptr;
    fb_fv *e = (fb_fv *)ext(env, a.v[0]);
    if (!e) return nullptr;
    int32_t n = 0;
    FB_OK_OR_THROW(fb_fv_tape_len(e, &n), "fvTapeLen");
    return make_i32(env, n);
}

// fvBackward(e, dLdV, gWeight, gBias, gKdt, dLdLum | null): accumulators added to in place
napi_value fv_backward(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 6)) return nullptr;
    fb_fv *e = (fb_fv *)ext(env, a.v[0]);
    float *dv, *gw, *gb, *gk, *dl;
    size_t n0, n1, n2, n3, n4;
    if (!e || !typed(env, a.v[1], &dv, &n0) || !typed(env, a.v[2], &gw, &n1) || !typed(env, a.v[3], &gb, &n2) ||
        !typed(env, a.v[4], &gk, &n3) || !typed(env, a.v[5], &dl, &n4, true))
        return nullptr;
    size_t N, E, nCol, tape;
    if (!fv_dims(env, e, &N, &E, &nCol, &tape) || !long_enough(env, n0, tape * N, "fvBackward dLdV") ||
        !long_enough(env, n1, E, "fvBackward gWeight") || !long_enough(env, n2, N, "fvBackward gBias") ||
        !long_enough(env, n3, N, "fvBackward gKdt") || (dl && !long_enough(env, n4, tape * nCol, "fvBackward dLdLum")))
        return nullptr;
    FB_OK_OR_THROW(fb_fv_backward(e, dv, gw, gb, gk, dl), "fvBackward");
    return nullptr;
}

// ---------------------------------------------------------------- differentiable model (fb_diff_*)
//
// Handles of the differentiable model, its lockstep batches and the ensemble
// engine.  Arrays are checked against the sizes the C side copies (fb_diff_size,
// fb_batch_size) before any copy, as for the brain.

std::string str_arg(napi_env env, napi_value v) {
    char buf[128] = {0};
    size_t n = 0;
    napi_get_value_string_utf8(env, v, buf, sizeof buf, &n);
    return std::string(buf, n);
}

bool is_nullish(napi_env env, napi_value v) {
    napi_valuetype t;
    napi_typeof(env, v, &t);
    return t == napi_null || t == napi_undefined;
}

// fb_diff_params from a Float64Array: the 23 doubles in field order, then
// soft, checkpoint and exact32
constexpr size_t DIFF_DOUBLES = 23, DIFF_PARAMS_LEN = DIFF_DOUBLES + 3;
bool read_diff_params(napi_env env, napi_value v, fb_diff_params *p) {
    double *d;
    size_t n;
    if (!typed(env, v, &d, &n)) return false;
    if (n != DIFF_PARAMS_LEN) {
        napi_throw_range_error(env, nullptr, "diff params: wrong length for fb_diff_params");
        return false;
    }
    std::memset(p, 0, sizeof *p);
    std::memcpy(p, d, sizeof(double) * DIFF_DOUBLES);
    p->soft = (int32_t)d[DIFF_DOUBLES];
    p->checkpoint = (int32_t)d[DIFF_DOUBLES + 1];
    p->exact32 = (int32_t)d[DIFF_DOUBLES + 2];
    return true;
}

struct DiffDims {
    size_t N, E, steps, R;
};

bool diff_dims(napi_env env, fb_diff *f, DiffDims *dm) {
    int32_t N = 0, T = 0, R = 0;
    int64_t E = 0;
    fb_status st = fb_diff_size(f, &N, &E, &T, &R);
    if (st != FB_OK) { throw_fb(env, st, "diffSize"); return false; }
    dm->N = (size_t)N; dm->E = (size_t)E; dm->steps = (size_t)T; dm->R = (size_t)R;
    return true;
}

#define DIFF_OR_NULL(idx)                         \
    fb_diff *f = (fb_diff *)ext(env, a.v[idx]);   \
    if (!f) return nullptr;                       \
    DiffDims dm;                                  \
    if (!diff_dims(env, f, &dm)) return nullptr

napi_value diff_params_length(napi_env env, napi_callback_info) {
    return make_i32(env, (int32_t)DIFF_PARAMS_LEN);
}

// diffCreate(N, Uint32Array indptr, Uint32Array indices, Float32Array counts, Float32Array sign, Float64Array params)
napi_value diff_create(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 6)) return nullptr;
    const int32_t N = i32(env, a.v[0]);
    uint32_t *ip, *ix;
    float *cn, *sg;
    size_t nip, nix, nc, ns;
    if (!typed(env, a.v[1], &ip, &nip) || !typed(env, a.v[2], &ix, &nix) || !typed(env, a.v[3], &cn, &nc) ||
        !typed(env, a.v[4], &sg, &ns))
        return nullptr;
    if (N <= 0 || nip != (size_t)N + 1 || nc != nix || ns != (size_t)N) {
        napi_throw_range_error(env, nullptr, "diffCreate: array sizes do not match N and E");
        return nullptr;
    }
    fb_diff_params p;
    if (!read_diff_params(env, a.v[5], &p)) return nullptr;
    fb_diff *f = nullptr;
    FB_OK_OR_THROW(fb_diff_create(N, (int64_t)nix, ip, ix, cn, sg, &p, &f), "diffCreate");
    return make_ext(env, f);
}

napi_value diff_destroy(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 1)) return nullptr;
    fb_diff *f = (fb_diff *)ext(env, a.v[0]);
    if (!f) return nullptr;
    FB_OK_OR_THROW(fb_diff_destroy(f), "diffDestroy");
    return nullptr;
}

napi_value diff_set_params(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 2)) return nullptr;
    DIFF_OR_NULL(0);
    fb_diff_params p;
    if (!read_diff_params(env, a.v[1], &p)) return nullptr;
    FB_OK_OR_THROW(fb_diff_set_params(f, &p), "diffSetParams");
    return nullptr;
}

// diffSetInput(h, which, Float32Array | null)
napi_value diff_set_input(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 3)) return nullptr;
    DIFF_OR_NULL(0);
    float *x;
    size_t n;
    if (!typed(env, a.v[2], &x, &n, true)) return nullptr;
    if (x && !long_enough(env, n, dm.N, "diffSetInput")) return nullptr;
    FB_OK_OR_THROW(fb_diff_set_input(f, i32(env, a.v[1]), x), "diffSetInput");
    return nullptr;
}

napi_value diff_set_edge_gain(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 2)) return nullptr;
    DIFF_OR_NULL(0);
    float *x;
    size_t n;
    if (!typed(env, a.v[1], &x, &n, true)) return nullptr;
    if (x && !long_enough(env, n, dm.E, "diffSetEdgeGain")) return nullptr;
    FB_OK_OR_THROW(fb_diff_set_edge_gain(f, x), "diffSetEdgeGain");
    return nullptr;
}

// diffSetMasks(h, Uint8Array thrMask | null, Uint8Array biasMask | null)
napi_value diff_set_masks(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 3)) return nullptr;
    DIFF_OR_NULL(0);
    uint8_t *t, *b;
    size_t nt, nb;
    if (!typed(env, a.v[1], &t, &nt, true) || !typed(env, a.v[2], &b, &nb, true)) return nullptr;
    if ((t && !long_enough(env, nt, dm.N, "diffSetMasks thrMask")) ||
        (b && !long_enough(env, nb, dm.N, "diffSetMasks biasMask")))
        return nullptr;
    FB_OK_OR_THROW(fb_diff_set_masks(f, t, b), "diffSetMasks");
    return nullptr;
}

napi_value diff_set_driven(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 2)) return nullptr;
    DIFF_OR_NULL(0);
    int32_t *ix;
    size_t n;
    if (!typed(env, a.v[1], &ix, &n)) return nullptr;
    FB_OK_OR_THROW(fb_diff_set_driven(f, ix, (int32_t)n), "diffSetDriven");
    return nullptr;
}

// diffSetLoss(h, Float32Array dLdSpike | null, lossFrom)
napi_value diff_set_loss(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 3)) return nullptr;
    DIFF_OR_NULL(0);
    float *x;
    size_t n;
    if (!typed(env, a.v[1], &x, &n, true)) return nullptr;
    if (x && !long_enough(env, n, dm.N, "diffSetLoss")) return nullptr;
    FB_OK_OR_THROW(fb_diff_set_loss(f, x, i32(env, a.v[2])), "diffSetLoss");
    return nullptr;
}

// diffSetReadout(h, Int32Array idx, Float32Array targetHz, tauMs)
napi_value diff_set_readout(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 4)) return nullptr;
    DIFF_OR_NULL(0);
    int32_t *ix;
    float *tg;
    size_t n, nt;
    if (!typed(env, a.v[1], &ix, &n) || !typed(env, a.v[2], &tg, &nt)) return nullptr;
    if (!long_enough(env, nt, n, "diffSetReadout targets")) return nullptr;
    FB_OK_OR_THROW(fb_diff_set_readout(f, ix, (int32_t)n, tg, num(env, a.v[3])), "diffSetReadout");
    return nullptr;
}

napi_value diff_set_tape_capacity(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 2)) return nullptr;
    DIFF_OR_NULL(0);
    FB_OK_OR_THROW(fb_diff_set_tape_capacity(f, (int64_t)num(env, a.v[1])), "diffSetTapeCapacity");
    return nullptr;
}

// diffSetDriveSchedule(h, Int32Array idx | null, epochSteps, Float32Array sched, nEpochs, driveSoft)
napi_value diff_set_drive_schedule(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 6)) return nullptr;
    DIFF_OR_NULL(0);
    int32_t *ix;
    float *sc;
    size_t n, ns;
    if (!typed(env, a.v[1], &ix, &n, true) || !typed(env, a.v[3], &sc, &ns, true)) return nullptr;
    const int32_t nEp = i32(env, a.v[4]);
    if (ix && (nEp < 1 || !sc || !long_enough(env, ns, n * (size_t)nEp, "diffSetDriveSchedule"))) {
        if (!sc || nEp < 1) napi_throw_range_error(env, nullptr, "diffSetDriveSchedule: no schedule");
        return nullptr;
    }
    FB_OK_OR_THROW(fb_diff_set_drive_schedule(f, ix, ix ? (int32_t)n : 0, i32(env, a.v[2]), sc, nEp,
                                              i32(env, a.v[5])),
                   "diffSetDriveSchedule");
    return nullptr;
}

// diffSetLossPerEpoch(h, Float32Array dPer | null, nEpochs, epochSteps)
napi_value diff_set_loss_per_epoch(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 4)) return nullptr;
    DIFF_OR_NULL(0);
    float *x;
    size_t n;
    if (!typed(env, a.v[1], &x, &n, true)) return nullptr;
    const int32_t nEp = i32(env, a.v[2]);
    if (x && (nEp < 1 || !long_enough(env, n, (size_t)nEp * dm.N, "diffSetLossPerEpoch"))) {
        if (nEp < 1) napi_throw_range_error(env, nullptr, "diffSetLossPerEpoch: nEpochs < 1");
        return nullptr;
    }
    FB_OK_OR_THROW(fb_diff_set_loss_per_epoch(f, x, nEp, i32(env, a.v[3])), "diffSetLossPerEpoch");
    return nullptr;
}

// diffDriveGrad(h, Float32Array out, length): out holds at least `length`
// values, the schedule's nEpochs * nD
napi_value diff_drive_grad(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 3)) return nullptr;
    DIFF_OR_NULL(0);
    float *x;
    size_t n;
    if (!typed(env, a.v[1], &x, &n) || !long_enough(env, n, (size_t)num(env, a.v[2]), "diffDriveGrad")) return nullptr;
    FB_OK_OR_THROW(fb_diff_drive_grad(f, x), "diffDriveGrad");
    return nullptr;
}

// diffSetSpikeTrain(h, Float64Array ptr | null, Int32Array idx | null): step t
// forces idx[ptr[t] .. ptr[t+1]) (ptr holds steps + 1 integral offsets from
// 0); null clears
napi_value diff_set_spike_train(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 3)) return nullptr;
    DIFF_OR_NULL(0);
    double *pd;
    int32_t *ix;
    size_t np_ = 0, ni = 0;
    if (!typed(env, a.v[1], &pd, &np_, true) || !typed(env, a.v[2], &ix, &ni, true)) return nullptr;
    if (!pd || np_ < 2) {
        FB_OK_OR_THROW(fb_diff_set_spike_train(f, 0, nullptr, nullptr), "diffSetSpikeTrain");
        return nullptr;
    }
    std::vector<int64_t> ptr(np_);
    for (size_t k = 0; k < np_; k++) ptr[k] = (int64_t)pd[k];
    if (ptr[np_ - 1] < 0 || !long_enough(env, ni, (size_t)ptr[np_ - 1], "diffSetSpikeTrain")) {
        if (ptr[np_ - 1] < 0) napi_throw_range_error(env, nullptr, "diffSetSpikeTrain: negative offset");
        return nullptr;
    }
    FB_OK_OR_THROW(fb_diff_set_spike_train(f, (int32_t)(np_ - 1), ptr.data(), ix), "diffSetSpikeTrain");
    return nullptr;
}

// diffForward(h, steps, seed, record) -> loss
napi_value diff_forward(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 4)) return nullptr;
    DIFF_OR_NULL(0);
    double L = 0;
    FB_OK_OR_THROW(fb_diff_forward(f, i32(env, a.v[1]), num(env, a.v[2]), i32(env, a.v[3]), &L), "diffForward");
    return make_num(env, L);
}

// diffBackward(h, truncate, Float64Array params[9], Float32Array gLogGain[N],
//   Float32Array gEdge[E] | null, Float32Array gInScale[N], Int32Array clamped[3])
// truncate < 0: read the results of the last backward (a batch's) instead
napi_value diff_backward(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 7)) return nullptr;
    DIFF_OR_NULL(0);
    double *pp;
    float *lg, *ge, *gi;
    int32_t *cl;
    size_t n0, n1, n2, n3, n4;
    if (!typed(env, a.v[2], &pp, &n0) || !typed(env, a.v[3], &lg, &n1) || !typed(env, a.v[4], &ge, &n2, true) ||
        !typed(env, a.v[5], &gi, &n3) || !typed(env, a.v[6], &cl, &n4))
        return nullptr;
    if (!long_enough(env, n0, FB_G_NPARAMS, "diffBackward params") || !long_enough(env, n1, dm.N, "diffBackward logGain") ||
        (ge && !long_enough(env, n2, dm.E, "diffBackward edge")) || !long_enough(env, n3, dm.N, "diffBackward inScale") ||
        !long_enough(env, n4, 3, "diffBackward clamped"))
        return nullptr;
    const int32_t tr = i32(env, a.v[1]);
    if (tr < 0) FB_OK_OR_THROW(fb_diff_grads(f, pp, lg, ge, gi, cl), "diffGrads");
    else FB_OK_OR_THROW(fb_diff_backward(f, tr, pp, lg, ge, gi, cl), "diffBackward");
    return nullptr;
}

// diffTypeGrads(h, Int32Array typeOf[N], nTypes, Float64Array out[nTypes])
napi_value diff_type_grads(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 4)) return nullptr;
    DIFF_OR_NULL(0);
    int32_t *ty;
    double *out;
    size_t n, no;
    const int32_t nT = i32(env, a.v[2]);
    if (!typed(env, a.v[1], &ty, &n) || !typed(env, a.v[3], &out, &no)) return nullptr;
    if (nT < 1 || !long_enough(env, n, dm.N, "diffTypeGrads typeOf") || !long_enough(env, no, (size_t)nT, "diffTypeGrads out")) {
        if (nT < 1) napi_throw_range_error(env, nullptr, "diffTypeGrads: nTypes < 1");
        return nullptr;
    }
    FB_OK_OR_THROW(fb_diff_type_grads(f, ty, nT, out), "diffTypeGrads");
    return nullptr;
}

napi_value diff_get_state(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 3)) return nullptr;
    DIFF_OR_NULL(0);
    float *x;
    size_t n;
    if (!typed(env, a.v[2], &x, &n) || !long_enough(env, n, dm.N, "diffGetState")) return nullptr;
    FB_OK_OR_THROW(fb_diff_get_state(f, i32(env, a.v[1]), x), "diffGetState");
    return nullptr;
}

// diffSpikes(h, Float64Array off[steps+1] | null, Int32Array idx | null, Float32Array amp | null) -> total
napi_value diff_spikes(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 4)) return nullptr;
    DIFF_OR_NULL(0);
    double *off;
    int32_t *ix;
    float *am;
    size_t no, ni, na;
    if (!typed(env, a.v[1], &off, &no, true) || !typed(env, a.v[2], &ix, &ni, true) || !typed(env, a.v[3], &am, &na, true))
        return nullptr;
    if (off && !long_enough(env, no, dm.steps + 1, "diffSpikes offsets")) return nullptr;
    std::vector<int64_t> o(dm.steps + 1);
    const int64_t cap = (ix && am) ? (int64_t)std::min(ni, na) : 0;
    int64_t total = 0;
    FB_OK_OR_THROW(fb_diff_spikes(f, o.data(), cap ? ix : nullptr, cap ? am : nullptr, cap, &total), "diffSpikes");
    if (off)
        for (size_t k = 0; k <= dm.steps; k++) off[k] = (double)o[k];
    return make_num(env, (double)total);
}

napi_value diff_rng(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 1)) return nullptr;
    DIFF_OR_NULL(0);
    uint32_t r = 0;
    FB_OK_OR_THROW(fb_diff_rng(f, &r), "diffRng");
    return make_num(env, (double)r);
}

napi_value diff_stat(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 2)) return nullptr;
    DIFF_OR_NULL(0);
    double v = 0;
    FB_OK_OR_THROW(fb_diff_stat(f, str_arg(env, a.v[1]).c_str(), &v), "diffStat");
    return make_num(env, v);
}

// diffSize(h) -> [N, E, steps, R]
napi_value diff_size(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 1)) return nullptr;
    DIFF_OR_NULL(0);
    napi_value arr;
    NAPI_OK(napi_create_array_with_length(env, 4, &arr));
    const size_t v[4] = {dm.N, dm.E, dm.steps, dm.R};
    for (int k = 0; k < 4; k++) napi_set_element(env, arr, k, make_num(env, (double)v[k]));
    return arr;
}

// diffBatchCreate([handles]) -> handle
napi_value diff_batch_create(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 1)) return nullptr;
    uint32_t n = 0;
    bool isArr = false;
    napi_is_array(env, a.v[0], &isArr);
    if (!isArr || napi_get_array_length(env, a.v[0], &n) != napi_ok || n == 0) {
        napi_throw_type_error(env, nullptr, "diffBatchCreate: expected a non-empty array of handles");
        return nullptr;
    }
    std::vector<fb_diff *> ms(n);
    for (uint32_t k = 0; k < n; k++) {
        napi_value e;
        NAPI_OK(napi_get_element(env, a.v[0], k, &e));
        ms[k] = (fb_diff *)ext(env, e);
        if (!ms[k]) return nullptr;
    }
    fb_diff_batch *bt = nullptr;
    FB_OK_OR_THROW(fb_diff_batch_create(ms.data(), (int32_t)n, &bt), "diffBatchCreate");
    return make_ext(env, bt);
}

napi_value diff_batch_destroy(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 1)) return nullptr;
    fb_diff_batch *bt = (fb_diff_batch *)ext(env, a.v[0]);
    if (!bt) return nullptr;
    FB_OK_OR_THROW(fb_diff_batch_destroy(bt), "diffBatchDestroy");
    return nullptr;
}

// diffBatchForward(h, Int32Array steps[B], Float64Array seeds[B], record, Float64Array losses[B])
napi_value diff_batch_forward(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 5)) return nullptr;
    fb_diff_batch *bt = (fb_diff_batch *)ext(env, a.v[0]);
    int32_t *st;
    double *sd, *ls;
    size_t n0, n1, n2;
    if (!bt || !typed(env, a.v[1], &st, &n0) || !typed(env, a.v[2], &sd, &n1) || !typed(env, a.v[4], &ls, &n2))
        return nullptr;
    double B = 0;
    FB_OK_OR_THROW(fb_diff_batch_stat(bt, "members", &B), "diffBatchForward");
    if (!long_enough(env, n0, (size_t)B, "diffBatchForward steps") ||
        !long_enough(env, n1, (size_t)B, "diffBatchForward seeds") || !long_enough(env, n2, (size_t)B, "diffBatchForward losses"))
        return nullptr;
    FB_OK_OR_THROW(fb_diff_batch_forward(bt, st, sd, i32(env, a.v[3]), ls), "diffBatchForward");
    return nullptr;
}

napi_value diff_batch_backward(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 2)) return nullptr;
    fb_diff_batch *bt = (fb_diff_batch *)ext(env, a.v[0]);
    if (!bt) return nullptr;
    FB_OK_OR_THROW(fb_diff_batch_backward(bt, i32(env, a.v[1])), "diffBatchBackward");
    return nullptr;
}

napi_value diff_batch_stat(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 2)) return nullptr;
    fb_diff_batch *bt = (fb_diff_batch *)ext(env, a.v[0]);
    if (!bt) return nullptr;
    double v = 0;
    FB_OK_OR_THROW(fb_diff_batch_stat(bt, str_arg(env, a.v[1]).c_str(), &v), "diffBatchStat");
    return make_num(env, v);
}

// ---------------------------------------------------------------- ensembles (fb_batch_*)

struct BatchDims {
    size_t N, B, T;
};

bool batch_dims(napi_env env, fb_batch *b, BatchDims *dm) {
    int32_t N = 0, B = 0, T = 0;
    fb_status st = fb_batch_size(b, &N, &B, &T);
    if (st != FB_OK) { throw_fb(env, st, "batchSize"); return false; }
    dm->N = (size_t)N; dm->B = (size_t)B; dm->T = (size_t)T;
    return true;
}

#define BATCH_OR_NULL(idx)                          \
    fb_batch *bt = (fb_batch *)ext(env, a.v[idx]);  \
    if (!bt) return nullptr;                        \
    BatchDims dm;                                   \
    if (!batch_dims(env, bt, &dm)) return nullptr

// batchCreate(graph, Float64Array params, Float64Array seeds[B], Int32Array typeOf[N], nTypes, Float32Array gains[B*nTypes])
napi_value batch_create(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 6)) return nullptr;
    fb_graph *g = (fb_graph *)ext(env, a.v[0]);
    if (!g) return nullptr;
    fb_params p;
    if (!read_params(env, a.v[1], &p)) return nullptr;
    double *sd;
    int32_t *ty;
    float *gn;
    size_t nsd, nty, ngn;
    if (!typed(env, a.v[2], &sd, &nsd) || !typed(env, a.v[3], &ty, &nty) || !typed(env, a.v[5], &gn, &ngn)) return nullptr;
    const int32_t nT = i32(env, a.v[4]);
    if (nT < 1 || ngn < nsd * (size_t)nT) {
        napi_throw_range_error(env, nullptr, "batchCreate: gains must hold B * nTypes values");
        return nullptr;
    }
    fb_batch *bt = nullptr;
    FB_OK_OR_THROW(fb_batch_create(g, &p, (int32_t)nsd, sd, ty, nT, gn, &bt), "batchCreate");
    BatchDims dm;
    if (!batch_dims(env, bt, &dm)) return nullptr;
    if (nty < dm.N) {
        fb_batch_destroy(bt);
        napi_throw_range_error(env, nullptr, "batchCreate: typeOf must hold N values");
        return nullptr;
    }
    return make_ext(env, bt);
}

napi_value batch_destroy(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 1)) return nullptr;
    fb_batch *bt = (fb_batch *)ext(env, a.v[0]);
    if (!bt) return nullptr;
    FB_OK_OR_THROW(fb_batch_destroy(bt), "batchDestroy");
    return nullptr;
}

napi_value batch_set_drive(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 3)) return nullptr;
    BATCH_OR_NULL(0);
    int32_t *ix;
    size_t n;
    if (!typed(env, a.v[1], &ix, &n)) return nullptr;
    FB_OK_OR_THROW(fb_batch_set_drive(bt, ix, (int32_t)n, num(env, a.v[2])), "batchSetDrive");
    return nullptr;
}

// batchSetMember(h, which, member, Int32Array idx, value)
napi_value batch_set_member(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 5)) return nullptr;
    BATCH_OR_NULL(0);
    int32_t *ix;
    size_t n;
    if (!typed(env, a.v[3], &ix, &n)) return nullptr;
    FB_OK_OR_THROW(fb_batch_set_member(bt, i32(env, a.v[1]), i32(env, a.v[2]), ix, (int32_t)n, num(env, a.v[4])),
                   "batchSetMember");
    return nullptr;
}

napi_value batch_set_steps(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 2)) return nullptr;
    BATCH_OR_NULL(0);
    int32_t *st;
    size_t n;
    if (!typed(env, a.v[1], &st, &n) || !long_enough(env, n, dm.B, "batchSetSteps")) return nullptr;
    FB_OK_OR_THROW(fb_batch_set_steps(bt, st), "batchSetSteps");
    return nullptr;
}

// batchRun(h, nsteps)
napi_value batch_run(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 2)) return nullptr;
    BATCH_OR_NULL(0);
    FB_OK_OR_THROW(fb_batch_run(bt, i32(env, a.v[1]), nullptr, nullptr, nullptr, nullptr, 0, nullptr), "batchRun");
    return nullptr;
}

// batchGetArray(h, which, member, TypedArray out[N])
napi_value batch_get_array(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 4)) return nullptr;
    BATCH_OR_NULL(0);
    float *x;
    size_t n;
    if (!typed(env, a.v[3], &x, &n) || !long_enough(env, n, dm.N, "batchGetArray")) return nullptr;
    FB_OK_OR_THROW(fb_batch_get_array(bt, i32(env, a.v[1]), i32(env, a.v[2]), x), "batchGetArray");
    return nullptr;
}

// batchSetLanes(h, signGain, Float32Array laneDrive[B*N] | null)
napi_value batch_set_lanes(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 3)) return nullptr;
    BATCH_OR_NULL(0);
    float *x;
    size_t n;
    if (!typed(env, a.v[2], &x, &n, true)) return nullptr;
    if (x && !long_enough(env, n, dm.B * dm.N, "batchSetLanes")) return nullptr;
    FB_OK_OR_THROW(fb_batch_set_lanes(bt, i32(env, a.v[1]), x), "batchSetLanes");
    return nullptr;
}

napi_value batch_set_rng(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 2)) return nullptr;
    BATCH_OR_NULL(0);
    uint32_t *x;
    size_t n;
    if (!typed(env, a.v[1], &x, &n) || !long_enough(env, n, dm.B, "batchSetRng")) return nullptr;
    FB_OK_OR_THROW(fb_batch_set_rng(bt, x), "batchSetRng");
    return nullptr;
}

// batchSetPoolFeedback(h, Uint8Array poolOf[N] | null, nPools, Int32Array targets, base, gain, Int32Array delay[B], recordCap)
napi_value batch_set_pool_feedback(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 8)) return nullptr;
    BATCH_OR_NULL(0);
    uint8_t *po;
    int32_t *tg, *dl;
    size_t np, nt, nd;
    if (!typed(env, a.v[1], &po, &np, true) || !typed(env, a.v[3], &tg, &nt) || !typed(env, a.v[6], &dl, &nd)) return nullptr;
    if ((po && !long_enough(env, np, dm.N, "batchSetPoolFeedback poolOf")) ||
        (po && !long_enough(env, nd, dm.B, "batchSetPoolFeedback delay")))
        return nullptr;
    FB_OK_OR_THROW(fb_batch_set_pool_feedback(bt, po, i32(env, a.v[2]), tg, (int32_t)nt, (float)num(env, a.v[4]),
                                              (float)num(env, a.v[5]), dl, i32(env, a.v[7])),
                   "batchSetPoolFeedback");
    return nullptr;
}

// batchPoolCounts(h, Uint16Array dst | null, cap) -> records in the loop's log
napi_value batch_pool_counts(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 3)) return nullptr;
    BATCH_OR_NULL(0);
    uint16_t *x;
    size_t n;
    if (!typed(env, a.v[1], &x, &n, true)) return nullptr;
    int32_t nr = 0;
    const int32_t cap = x ? i32(env, a.v[2]) : 0;
    if (x) {
        // dst [B][min(cap, records)][pools]
        double pools = 0;
        FB_OK_OR_THROW(fb_batch_stat(bt, "pools", &pools), "batchPoolCounts");
        FB_OK_OR_THROW(fb_batch_pool_counts(bt, nullptr, 0, &nr), "batchPoolCounts");
        if (cap < 0 || !long_enough(env, n, dm.B * (size_t)std::min(cap, nr) * (size_t)pools, "batchPoolCounts"))
            return nullptr;
    }
    FB_OK_OR_THROW(fb_batch_pool_counts(bt, x, cap, &nr), "batchPoolCounts");
    return make_i32(env, nr);
}

napi_value batch_rng(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 2)) return nullptr;
    BATCH_OR_NULL(0);
    uint32_t *x;
    size_t n;
    if (!typed(env, a.v[1], &x, &n) || !long_enough(env, n, dm.B, "batchRng")) return nullptr;
    FB_OK_OR_THROW(fb_batch_rng(bt, x), "batchRng");
    return nullptr;
}

napi_value batch_counters(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 1)) return nullptr;
    BATCH_OR_NULL(0);
    int64_t c[4] = {0, 0, 0, 0};
    FB_OK_OR_THROW(fb_batch_counters(bt, &c[0], &c[1], &c[2], &c[3]), "batchCounters");
    napi_value arr;
    NAPI_OK(napi_create_array_with_length(env, 4, &arr));
    for (int k = 0; k < 4; k++) napi_set_element(env, arr, k, make_num(env, (double)c[k]));
    return arr;
}

napi_value batch_set_engine(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 2)) return nullptr;
    BATCH_OR_NULL(0);
    FB_OK_OR_THROW(fb_batch_set_engine(bt, i32(env, a.v[1])), "batchSetEngine");
    return nullptr;
}

napi_value batch_stat(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 2)) return nullptr;
    BATCH_OR_NULL(0);
    double v = 0;
    FB_OK_OR_THROW(fb_batch_stat(bt, str_arg(env, a.v[1]).c_str(), &v), "batchStat");
    return make_num(env, v);
}

// batchSize(h) -> [N, B, nTypes]
napi_value batch_size(napi_env env, napi_callback_info info) {
    Args a;
    if (!get_args(env, info, a, 1)) return nullptr;
    BATCH_OR_NULL(0);
    napi_value arr;
    NAPI_OK(napi_create_array_with_length(env, 3, &arr));
    const size_t v[3] = {dm.N, dm.B, dm.T};
    for (int k = 0; k < 3; k++) napi_set_element(env, arr, k, make_num(env, (double)v[k]));
    return arr;
}

napi_value init(napi_env env, napi_value exports) {
    struct Fn {
        const char *name;
        napi_callback cb;
    } fns[] = {
        {"graphCreate", graph_create},
        {"graphRelease", graph_release},
        {"brainCreate", brain_create},
        {"brainDestroy", brain_destroy},
        {"reset", reset},
        {"setParams", set_params},
        {"setDrive", list_setter<fb_set_drive>},
        {"setDriveOne", set_drive_one},
        {"setBias", list_setter<fb_set_bias>},
        {"setThr", set_thr},
        {"addG", add_g},
        {"applyDeltas", apply_deltas},
        {"pulse", list_setter<fb_pulse>},
        {"setBackground", set_background},
        {"step", step},
        {"fired", fired},
        {"getArray", get_array},
        {"setArray", set_array},
        {"scalars", scalars},
        {"counters", counters},
        {"runAsync", run_async},
        {"deviceInfo", device_info},
        {"paramsLength", params_length},
        {"fvModelCreate", fv_model_create},
        {"fvModelSetKdt", fv_model_set_kdt},
        {"fvModelSetInputMap", fv_model_set_input_map},
        {"fvModelRelease", fv_model_release},
        {"fvCreate", fv_create},
        {"fvDestroy", fv_call0<fb_fv_destroy>},
        {"fvReset", fv_call0<fb_fv_reset>},
        {"fvSetInput", fv_call_arr<fb_fv_set_input, false>},
        {"fvSetLum", fv_call_arr<fb_fv_set_lum, true>},
        {"fvSetV", fv_call_arr<fb_fv_set_v, false>},
        {"fvStep", fv_call_int<fb_fv_step>},
        {"fvTapeStart", fv_call_int<fb_fv_tape_start>},
        {"fvStepTaped", fv_call_int<fb_fv_step_taped>},
        {"fvTapeLen", fv_tape_len},
        {"fvGet", fv_get},
        {"fvBackward", fv_backward},
        {"diffParamsLength", diff_params_length},
        {"diffCreate", diff_create},
        {"diffDestroy", diff_destroy},
        {"diffSetParams", diff_set_params},
        {"diffSetInput", diff_set_input},
        {"diffSetEdgeGain", diff_set_edge_gain},
        {"diffSetMasks", diff_set_masks},
        {"diffSetDriven", diff_set_driven},
        {"diffSetLoss", diff_set_loss},
        {"diffSetReadout", diff_set_readout},
        {"diffSetTapeCapacity", diff_set_tape_capacity},
        {"diffSetDriveSchedule", diff_set_drive_schedule},
        {"diffSetLossPerEpoch", diff_set_loss_per_epoch},
        {"diffDriveGrad", diff_drive_grad},
        {"diffSetSpikeTrain", diff_set_spike_train},
        {"diffForward", diff_forward},
        {"diffBackward", diff_backward},
        {"diffTypeGrads", diff_type_grads},
        {"diffGetState", diff_get_state},
        {"diffSpikes", diff_spikes},
        {"diffRng", diff_rng},
        {"diffStat", diff_stat},
        {"diffSize", diff_size},
        {"diffBatchCreate", diff_batch_create},
        {"diffBatchDestroy", diff_batch_destroy},
        {"diffBatchForward", diff_batch_forward},
        {"diffBatchBackward", diff_batch_backward},
        {"diffBatchStat", diff_batch_stat},
        {"batchCreate", batch_create},
        {"batchDestroy", batch_destroy},
        {"batchSetDrive", batch_set_drive},
        {"batchSetMember", batch_set_member},
        {"batchSetSteps", batch_set_steps},
        {"batchRun", batch_run},
        {"batchGetArray", batch_get_array},
        {"batchSetLanes", batch_set_lanes},
        {"batchSetRng", batch_set_rng},
        {"batchSetPoolFeedback", batch_set_pool_feedback},
        {"batchPoolCounts", batch_pool_counts},
        {"batchRng", batch_rng},
        {"batchCounters", batch_counters},
        {"batchSetEngine", batch_set_engine},
        {"batchStat", batch_stat},
        {"batchSize", batch_size},
    };
    for (const Fn &f : fns) {
        napi_value v;
        NAPI_OK(napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb, nullptr, &v));
        NAPI_OK(napi_set_named_property(env, exports, f.name, v));
    }
    return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
