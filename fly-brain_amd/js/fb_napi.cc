// fb_napi.cc -- N-API binding of libflybrain for Node (the reference's host
// runtime).  Functions map one-to-one onto include/flybrain.h; handles are
// napi externals, arrays are TypedArrays the caller owns (copied in or out,
// never retained).  Errors become JS exceptions carrying fb_last_error().
// lifhip.js builds the LIFWasm-compatible class on top of these functions,
// flyviship.js the FlyVis / FlyVisDiff classes on the fv* ones.
#include <node_api.h>

#include <cstdint>
#include <cstring>
#include <string>

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
