// fb_api.hip -- the C ABI (include/flybrain.h): graphs, brains, setters,
// stepping engines and state transfer.
//
// Host objects mirror the reference's JS wrappers: fb_brain holds what LIFWasm
// holds on the JS side (params, the insertion-ordered drivenSet, t, ring head;
// src/lifwasm.js:36-87) while the state arrays live on the device.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <numeric>

#include "fb_internal.h"

namespace fb {

static thread_local std::string g_err;

void set_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}

fb_status launch_step(const DevBrain &d, int32_t head, int32_t slot, bool serialDeliver, bool rowsUnique,
                      hipStream_t st);
fb_status launch_gather(const DevBrain &d, int32_t s0, int32_t count, int32_t *cnt, int32_t *dst, hipStream_t st);
fb_status persistent_supported(const DevBrain &d, bool serialDeliver, std::string *why);
uint32_t persistent_overflow_flags(const DevBrain &d);
fb_status launch_persistent(const DevBrain &d, int32_t head, int32_t nsteps, hipStream_t st, int64_t *fallbackSteps,
                            int64_t *launches, double *kernelMs, double *phaseUs, uint64_t drivenVersion);
void persistent_release(const int32_t *key);
void persistent_take_time(const int32_t *key, double *kernelMs);
// The lookahead engine leaves the driven neurons' state in its stage between
// launches: brings the brain's arrays up to date, and drops the stage's copy
// of what the caller changes next (state: the neurons' state arrays; params:
// bias, thr, drive, the driven list, the constants).
fb_status persistent_sync(const DevBrain &d, hipStream_t st, bool state, bool params);

// ------------------------------------------------------------------ jump tables

// Byte tables of M^(2^b), b = 0..31, where M is one xorshift32 step viewed as a
// 32x32 matrix over GF(2) (the map is linear: shifts and xors only).
static void build_jump_tables(std::vector<uint32_t> &tab) {
    auto step = [](uint32_t x) { x ^= x << 13; x ^= x >> 17; x ^= x << 5; return x; };
    uint32_t col[32], tmp[32];
    for (int b = 0; b < 32; b++) col[b] = step(1u << b);
    auto apply = [](const uint32_t *c, uint32_t x) {
        uint32_t r = 0;
        for (int b = 0; b < 32; b++) if ((x >> b) & 1u) r ^= c[b];
        return r;
    };
    tab.assign(32 * 1024, 0);
    for (int p = 0; p < 32; p++) {
        for (int byte = 0; byte < 4; byte++)
            for (int v = 0; v < 256; v++)
                tab[(size_t)p * 1024 + byte * 256 + v] = apply(col, (uint32_t)v << (8 * byte));
        for (int b = 0; b < 32; b++) tmp[b] = apply(col, col[b]);   // square
        memcpy(col, tmp, sizeof col);
    }
}

// Leapfrog matrices M^1..M^32, then M^(32h) for h = 0..63 (M^0 = identity),
// then M^2048,
// each as 8 nibble tables: entry [k][v] is the XOR of the columns 4k..4k+3
// selected by the bits of v.
static void put_nibbles(uint32_t *dst, const uint32_t *col) {
    for (int k = 0; k < 8; k++)
        for (int v = 0; v < 16; v++) {
            uint32_t r = 0;
            for (int i = 0; i < 4; i++)
                if ((v >> i) & 1) r ^= col[4 * k + i];
            dst[k * 16 + v] = r;
        }
}

static void build_leap_matrices(std::vector<uint32_t> &tab) {
    auto step = [](uint32_t x) { x ^= x << 13; x ^= x >> 17; x ^= x << 5; return x; };
    tab.assign((size_t)LEAP_TABLE_MATS * LEAP_WORDS, 0);
    uint32_t col[32];
    for (int b = 0; b < 32; b++) col[b] = 1u << b;   // M^0
    put_nibbles(&tab[(size_t)32 * LEAP_WORDS], col);
    for (int p = 1; p <= 2048; p++) {
        for (int b = 0; b < 32; b++) col[b] = step(col[b]);   // M^p
        if (p <= 32) put_nibbles(&tab[(size_t)(p - 1) * LEAP_WORDS], col);
        if (p % 32 == 0 && p / 32 < 64) put_nibbles(&tab[(size_t)(32 + p / 32) * LEAP_WORDS], col);
    }
    put_nibbles(&tab[(size_t)LEAP_FAR * LEAP_WORDS], col);   // M^2048
}

static std::mutex g_jump_mu;
static std::map<int, std::pair<uint32_t *, uint32_t *>> g_jump_dev;

static fb_status jump_table(uint32_t **out, uint32_t **leap) {
    int dev = 0;
    FB_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_jump_mu);
    auto it = g_jump_dev.find(dev);
    if (it != g_jump_dev.end()) { *out = it->second.first; *leap = it->second.second; return FB_OK; }
    std::vector<uint32_t> tab, lp;
    build_jump_tables(tab);
    build_leap_matrices(lp);
    uint32_t *p = nullptr, *q = nullptr;
    FB_HIP(hipMalloc(&p, tab.size() * sizeof(uint32_t)));
    FB_HIP(hipMemcpy(p, tab.data(), tab.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    FB_HIP(hipMalloc(&q, lp.size() * sizeof(uint32_t)));
    FB_HIP(hipMemcpy(q, lp.data(), lp.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    g_jump_dev[dev] = {p, q};
    *out = p;
    *leap = q;
    return FB_OK;
}

// ------------------------------------------------------------------ constants

static double js_round(double x) {
    double f = std::floor(x);
    return (x - f >= 0.5) ? f + 1.0 : f;
}

int32_t nslots_of(const fb_params &p) {
    double r = js_round(p.delay / p.dt);
    if (r < 1.0) r = 1.0;
    return (int32_t)r + 1;
}

void derive_consts(const fb_params &p, int32_t N, Consts &c) {
    c.dt = (float)p.dt; c.vRest = (float)p.vRest; c.vThresh = (float)p.vThresh; c.vReset = (float)p.vReset;
    c.tRef = (float)p.tRef; c.adaptInc = (float)p.adaptInc; c.depU = (float)p.depU;
    c.dE = (float)std::exp(-p.dt / p.tauSyn);
    c.dTr = (float)std::exp(-p.dt / p.traceTau);
    c.dA = (float)std::exp(-p.dt / p.adaptTau);
    c.kRec = (float)(p.dt / p.depTau);
    c.kM = (float)(p.dt / p.tauM);
    c.dtS = (float)(p.dt / 1000.0);
    c.eExc = (float)p.eExc; c.eInh = (float)p.eInh;
    c.cE = (float)(1.0 / (p.eExc - p.vRest));
    c.cI = (float)(1.0 / (p.vRest - p.eInh));
    c.bgEvents = (float)((double)N * p.bgRate * p.dt / 1000.0);
    c.bgAmp = (float)p.bgAmp;
    c.preInh = (float)p.preInh;
    c.coba = p.coba != 0.0 ? 1 : 0;
}

uint32_t seed_hash(double seed) {
    double prod = seed * 2654435761.0;
    double t = prod < 0 ? -std::floor(-prod) : std::floor(prod);
    double m = std::fmod(t, 4294967296.0);
    if (m < 0) m += 4294967296.0;
    uint32_t s = (uint32_t)m;
    return s ? s : 1u;
}

}  // namespace fb

using namespace fb;

// ------------------------------------------------------------------ objects

struct fb_graph {
    int refs = 1;
    int device = 0;
    DevGraph d{};
    bool rowsUnique = true;
    bool wNonneg = true;            // no weight carries a sign bit (negative or -0)
    std::vector<void *> allocs;
    std::vector<uint32_t> indptr;   // host copy for validation
    std::vector<uint64_t> chunkCost;   // DevGraph::chunkCost
};

struct GraphCache {
    int32_t head = -1;
    int32_t steps = 0;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
};

struct fb_brain {
    fb_graph *g = nullptr;
    fb_params p{};
    DevBrain d{};
    hipStream_t stream = nullptr;      // the stream all work is queued on
    hipStream_t ownStream = nullptr;   // the brain's own stream
    std::vector<int32_t> drivenList;   // insertion order of the driven set
    // No excitatory conductance can be -0: gE only ever received non-negative
    // additions and decays since the last reset.  With non-negative weights a
    // zero-weight arrival then changes nothing, and the deliveries may
    // skip it (an inhibitory one, adding -0 to gI, never changes anything).
    bool gENonneg = true;
    std::vector<uint8_t> inDriven;
    bool drivenDirty = true;
    int32_t head = 0, lastSlot = 0;
    double t = 0.0;
    int64_t steps = 0;
    int64_t fallbackSteps = 0, persistentLaunches = 0;
    double persistentMs = 0.0;   // summed HIP-event duration of the lookahead kernel launches
    double phaseUs[PERSIST_PROF] = {0};   // lookahead kernel clocks
    std::vector<void *> allocs;
    int32_t *ringAlloc = nullptr;
    int32_t *idxScratch = nullptr;    // device indices for multi-neuron setters
    void *deltaBuf = nullptr;         // fb_apply_deltas staging: values, indices, kinds, group offsets
    int64_t deltaCap = 0;
    int32_t *gatherCnt = nullptr, *gatherBuf = nullptr;
    std::map<std::pair<int32_t, int32_t>, GraphCache> graphs;   // keyed by (start head, steps)
    uint64_t version = 0;                   // bumps when DevBrain arguments change
    uint64_t drivenVersion = 0;             // bumps when the device driven list is rewritten
    uint64_t graphVersion = ~0ull;
};

extern "C" const char *fb_last_error(void) { return g_err.c_str(); }

namespace fb {
const DevGraph *graph_dev(const fb_graph *g) { return &g->d; }
fb_status leap_tables(uint32_t **jump, uint32_t **leap) { return jump_table(jump, leap); }
}  // namespace fb

#define FB_REQUIRE(cond, code, ...)      \
    do {                                 \
        if (!(cond)) {                   \
            set_error(__VA_ARGS__);      \
            return code;                 \
        }                                \
    } while (0)

template <class T>
static fb_status dalloc(std::vector<void *> &list, T **p, size_t count) {
    void *q = nullptr;
    FB_HIP(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
    list.push_back(q);
    *p = (T *)q;
    return FB_OK;
}

extern "C" fb_status fb_device_info(char *name, int32_t cap, int32_t *n_cu) {
    int dev = 0;
    FB_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    FB_HIP(hipGetDeviceProperties(&prop, dev));
    if (name && cap > 0) snprintf(name, (size_t)cap, "%s (%s)", prop.name, prop.gcnArchName);
    if (n_cu) *n_cu = prop.multiProcessorCount;
    return FB_OK;
}

// ------------------------------------------------------------------ graph

extern "C" fb_status fb_graph_create(int32_t N, int64_t E, const uint32_t *indptr, const uint32_t *indices,
                                     const float *weights, const float *sign, const int32_t *counts,
                                     fb_graph **out) {
    FB_REQUIRE(out && indptr && sign && N > 0 && E >= 0, FB_ERR_ARG, "fb_graph_create: bad arguments");
    FB_REQUIRE(E == 0 || (indices && weights), FB_ERR_ARG, "fb_graph_create: indices/weights missing");
    FB_REQUIRE(E < (int64_t)UINT32_MAX, FB_ERR_ARG, "fb_graph_create: E exceeds the u32 CSR layout");
    FB_REQUIRE(indptr[0] == 0 && (int64_t)indptr[N] == E, FB_ERR_ARG,
               "fb_graph_create: indptr must start at 0 and end at E");
    for (int32_t i = 0; i < N; i++)
        FB_REQUIRE(indptr[i] <= indptr[i + 1], FB_ERR_ARG, "fb_graph_create: indptr not monotone at %d", i);
    for (int64_t j = 0; j < E; j++)
        FB_REQUIRE(indices[j] < (uint32_t)N, FB_ERR_ARG, "fb_graph_create: target %u out of range at edge %lld",
                   indices[j], (long long)j);
    fb_graph *g = new fb_graph();
    FB_HIP(hipGetDevice(&g->device));
    g->indptr.assign(indptr, indptr + N + 1);
    // Rows are stored sorted by target (stable, so repeated targets keep their
    // relative order); the order key of an arrival is its position in this row.
    std::vector<uint2> edges((size_t)std::max<int64_t>(E, 1));
    std::vector<int32_t> cnts;
    if (counts) cnts.resize((size_t)E);
    std::vector<uint32_t> perm;
    for (int64_t j = 0; j < E; j++)
        if (std::signbit(weights[j])) g->wNonneg = false;
    for (int32_t i = 0; i < N; i++) {
        const uint32_t a = indptr[i], b = indptr[i + 1];
        bool sorted = true;
        for (uint32_t j = a + 1; j < b; j++) {
            if (indices[j] < indices[j - 1]) sorted = false;
            if (indices[j] == indices[j - 1]) g->rowsUnique = false;
        }
        if (sorted) {
            for (uint32_t j = a; j < b; j++) {
                uint32_t wb;
                memcpy(&wb, &weights[j], 4);
                edges[j] = make_uint2(indices[j], wb);
                if (counts) cnts[j] = counts[j];
            }
        } else {
            perm.resize(b - a);
            std::iota(perm.begin(), perm.end(), a);
            std::stable_sort(perm.begin(), perm.end(), [&](uint32_t x, uint32_t y) { return indices[x] < indices[y]; });
            for (uint32_t j = a; j < b; j++) {
                const uint32_t src = perm[j - a];
                uint32_t wb;
                memcpy(&wb, &weights[src], 4);
                edges[j] = make_uint2(indices[src], wb);
                if (counts) cnts[j] = counts[src];
            }
            for (uint32_t j = a + 1; j < b; j++)
                if (edges[j].x == edges[j - 1].x) g->rowsUnique = false;
        }
    }
    uint32_t *dIndptr; uint2 *dEdges; float *dSign; int32_t *dCounts = nullptr; uint32_t *dRowSyn, *dInOff;
    fb_status st;
    {
        // in-degree prefix: the per-step delivery keeps a target's later
        // same-step arrivals in entries inOff[q] .. inOff[q+1]-1
        std::vector<uint32_t> inOff((size_t)N + 1, 0);
        for (int64_t j = 0; j < E; j++) inOff[indices[j] + 1]++;
        for (int32_t q = 0; q < N; q++) inOff[q + 1] += inOff[q];
        if ((st = dalloc(g->allocs, &dInOff, (size_t)N + 1))) return st;
        FB_HIP(hipMemcpy(dInOff, inOff.data(), sizeof(uint32_t) * ((size_t)N + 1), hipMemcpyHostToDevice));
    }
    if ((st = dalloc(g->allocs, &dIndptr, (size_t)N + 1)) ||
        (st = dalloc(g->allocs, &dEdges, (size_t)std::max<int64_t>(E, 1))) ||
        (st = dalloc(g->allocs, &dSign, (size_t)N)) || (st = dalloc(g->allocs, &dRowSyn, (size_t)N)))
        return st;
    {
        std::vector<uint32_t> rs((size_t)N, 0);
        if (counts)
            for (int32_t i = 0; i < N; i++) {
                uint64_t s = 0;
                for (uint32_t j = indptr[i]; j < indptr[i + 1]; j++) s += (uint64_t)std::max(counts[j], 0);
                rs[i] = (uint32_t)std::min<uint64_t>(s, UINT32_MAX);
            }
        FB_HIP(hipMemcpy(dRowSyn, rs.data(), sizeof(uint32_t) * (size_t)N, hipMemcpyHostToDevice));
    }
    FB_HIP(hipMemcpy(dIndptr, indptr, sizeof(uint32_t) * ((size_t)N + 1), hipMemcpyHostToDevice));
    if (E) FB_HIP(hipMemcpy(dEdges, edges.data(), sizeof(uint2) * (size_t)E, hipMemcpyHostToDevice));
    FB_HIP(hipMemcpy(dSign, sign, sizeof(float) * (size_t)N, hipMemcpyHostToDevice));
    if (counts) {
        if ((st = dalloc(g->allocs, &dCounts, (size_t)std::max<int64_t>(E, 1)))) return st;
        if (E) FB_HIP(hipMemcpy(dCounts, cnts.data(), sizeof(int32_t) * (size_t)E, hipMemcpyHostToDevice));
    }
    for (int32_t i = 0; i < N; i++) g->d.maxRow = std::max<int64_t>(g->d.maxRow, (int64_t)indptr[i + 1] - indptr[i]);
    g->chunkCost.assign(((size_t)N + 63) / 64, 0);
    for (int32_t i = 0; i < N; i++) {
        g->chunkCost[(size_t)i >> 6] += 16;
        for (uint32_t j = indptr[i]; j < indptr[i + 1]; j++)
            if (weights[j] != 0.f) {
                g->chunkCost[(size_t)i >> 6]++;
                g->chunkCost[(size_t)indices[j] >> 6]++;
            }
    }
    g->d.chunkCost = g->chunkCost.data();
    g->d.N = N; g->d.E = E; g->d.indptr = dIndptr; g->d.edges = dEdges; g->d.sign = dSign; g->d.counts = dCounts; g->d.rowSyn = dRowSyn; g->d.inOff = dInOff;
    *out = g;
    return FB_OK;
}

extern "C" fb_status fb_graph_retain(fb_graph *g) {
    FB_REQUIRE(g, FB_ERR_ARG, "fb_graph_retain: null graph");
    g->refs++;
    return FB_OK;
}

extern "C" fb_status fb_graph_release(fb_graph *g) {
    FB_REQUIRE(g, FB_ERR_ARG, "fb_graph_release: null graph");
    if (--g->refs > 0) return FB_OK;
    for (void *p : g->allocs) hipFree(p);
    delete g;
    return FB_OK;
}

// ------------------------------------------------------------------ brain helpers

__global__ void k_reset_state(DevBrain d) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < d.N) {
        d.v[i] = d.k.vRest; d.gE[i] = 0.f; d.gI[i] = 0.f; d.refr[i] = 0.f;
        d.trace[i] = 0.f; d.adapt[i] = 0.f; d.res[i] = 1.f; d.spikeCount[i] = 0u;
    }
    if (i < d.nslots) d.ringCount[i] = 0;
}

__global__ void k_init_scratch(DevBrain d) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 2 * d.N) d.cnt[i] = 0;
    if (i < d.N) d.bgCnt[i] = 0;
}

// setters: mode 0 stores f32(value); mode 1 stores f32(f64(old) + value)
__global__ void k_set(float *arr, const int32_t *idx, int32_t n, int32_t one, double value, int32_t mode,
                      int32_t serial) {
    if (serial) {
        if (blockIdx.x || threadIdx.x) return;
        for (int32_t q = 0; q < n; q++) {
            const int32_t i = idx ? idx[q] : one;
            arr[i] = mode ? (float)((double)arr[i] + value) : (float)value;
        }
        return;
    }
    const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int32_t i = idx ? idx[q] : one;
    arr[i] = mode ? (float)((double)arr[i] + value) : (float)value;
}

static fb_status sync_driven(fb_brain *b) {
    if (!b->drivenDirty) return FB_OK;
    const size_t n = b->drivenList.size();
    if (n) FB_HIP(hipMemcpyAsync(b->d.driven, b->drivenList.data(), n * sizeof(int32_t), hipMemcpyHostToDevice,
                                 b->stream));
    FB_HIP(hipStreamSynchronize(b->stream));
    if (b->d.nDriven != (int32_t)n) b->version++;
    b->d.nDriven = (int32_t)n;
    b->drivenDirty = false;
    b->drivenVersion++;
    return FB_OK;
}

// Every entry point that reads or writes a brain's arrays, or changes what
// the lookahead engine stages, calls this first.
static fb_status sync_stage(fb_brain *b, bool state, bool params) {
    return persistent_sync(b->d, b->stream, state, params);
}

static float *array_ptr(fb_brain *b, int32_t which) {
    switch (which) {
        case FB_V: return b->d.v;
        case FB_GE: return b->d.gE;
        case FB_GI: return b->d.gI;
        case FB_REFR: return b->d.refr;
        case FB_TRACE: return b->d.trace;
        case FB_ADAPT: return b->d.adapt;
        case FB_RES: return b->d.res;
        case FB_BIAS: return b->d.bias;
        case FB_THR: return b->d.thr;
        case FB_DRIVE: return b->d.drive;
        case FB_SPIKECOUNT: return (float *)b->d.spikeCount;
        default: return nullptr;
    }
}

static fb_status alloc_ring(fb_brain *b) {
    const int32_t N = b->d.N;
    if (b->ringAlloc) {
        hipFree(b->ringAlloc);
        b->allocs.erase(std::find(b->allocs.begin(), b->allocs.end(), (void *)b->ringAlloc));
        b->ringAlloc = nullptr;
    }
    int32_t *ring;
    b->d.ringStride = 2 * (int64_t)N;
    fb_status st = dalloc(b->allocs, &ring, (size_t)b->d.nslots * (size_t)b->d.ringStride + (size_t)b->d.nslots);
    if (st) return st;
    b->ringAlloc = ring;
    b->d.ring = ring;
    b->d.ringCount = ring + (size_t)b->d.nslots * (size_t)b->d.ringStride;
    FB_HIP(hipMemsetAsync(b->d.ringCount, 0, sizeof(int32_t) * (size_t)b->d.nslots, b->stream));
    return FB_OK;
}

static void drop_graphs(fb_brain *b) {
    // an executable graph is destroyed only after its launches have finished:
    // the brain's last call may have left replays queued (a call that changes
    // the zero-weight skip flag drops the graphs before it queues anything)
    if (!b->graphs.empty()) (void)hipStreamSynchronize(b->stream);
    for (auto &kv : b->graphs) {
        if (kv.second.exec) hipGraphExecDestroy(kv.second.exec);
        if (kv.second.graph) hipGraphDestroy(kv.second.graph);
    }
    b->graphs.clear();
}

// ------------------------------------------------------------------ brain API

extern "C" fb_status fb_brain_create(fb_graph *g, const fb_params *p, double seed, fb_brain **out) {
    FB_REQUIRE(g && p && out, FB_ERR_ARG, "fb_brain_create: null argument");
    FB_REQUIRE(p->dt > 0 && p->tauM > 0 && p->tauSyn > 0, FB_ERR_ARG, "fb_brain_create: dt, tauM, tauSyn must be > 0");
    fb_brain *b = new fb_brain();
    b->g = g;
    g->refs++;
    b->p = *p;
    const int32_t N = g->d.N;
    const int64_t E = g->d.E;
    DevBrain &d = b->d;
    d.N = N;
    d.nslots = nslots_of(*p);
    d.nDriven = 0;
    d.nblocks = (N + 255) / 256;
    derive_consts(*p, N, d.k);
    d.g = g->d;
    FB_HIP(hipStreamCreateWithFlags(&b->ownStream, hipStreamNonBlocking));
    b->stream = b->ownStream;
    fb_status st;
    float *f;
    // ten f32 state arrays back to back, then spike counts
    if ((st = dalloc(b->allocs, &f, (size_t)N * 10))) return st;
    d.v = f; d.gE = f + N; d.gI = f + 2 * (size_t)N; d.refr = f + 3 * (size_t)N; d.trace = f + 4 * (size_t)N;
    d.adapt = f + 5 * (size_t)N; d.res = f + 6 * (size_t)N; d.bias = f + 7 * (size_t)N; d.thr = f + 8 * (size_t)N;
    d.drive = f + 9 * (size_t)N;
    if ((st = dalloc(b->allocs, &d.spikeCount, (size_t)N))) return st;
    if ((st = dalloc(b->allocs, &d.driven, (size_t)N))) return st;
    if ((st = dalloc(b->allocs, &d.cnt, 2 * (size_t)N))) return st;
    if ((st = dalloc(b->allocs, &d.add, 2 * (size_t)N))) return st;
    if ((st = dalloc(b->allocs, &d.key, 2 * (size_t)N))) return st;
    // later same-step arrivals: one run per (sign class, target) sized by its in-degree
    if ((st = dalloc(b->allocs, &d.runKey, 2 * (size_t)std::max<int64_t>(E, 1)))) return st;
    if ((st = dalloc(b->allocs, &d.runAdd, 2 * (size_t)std::max<int64_t>(E, 1)))) return st;
    if ((st = dalloc(b->allocs, &d.longList, 2 * (size_t)N))) return st;
    if ((st = dalloc(b->allocs, &d.bgCnt, (size_t)N))) return st;
    if ((st = dalloc(b->allocs, &d.stage, (size_t)d.nblocks * 256))) return st;
    if ((st = dalloc(b->allocs, &d.blkCnt, (size_t)d.nblocks))) return st;
    if ((st = dalloc(b->allocs, &d.sc, 1))) return st;
    if ((st = dalloc(b->allocs, &b->idxScratch, (size_t)N))) return st;
    if ((st = jump_table((uint32_t **)&d.jumpTab, (uint32_t **)&d.leap))) return st;
    if ((st = alloc_ring(b))) return st;
    Scalars sc{};
    sc.rng = seed_hash(seed);
    FB_HIP(hipMemcpyAsync(d.sc, &sc, sizeof sc, hipMemcpyHostToDevice, b->stream));
    FB_HIP(hipMemsetAsync(d.bias, 0, sizeof(float) * 3 * (size_t)N, b->stream));   // bias, thr, drive
    const int nb = (2 * N + 255) / 256;
    hipLaunchKernelGGL(k_init_scratch, dim3(nb), dim3(256), 0, b->stream, d);
    hipLaunchKernelGGL(k_reset_state, dim3((std::max(N, d.nslots) + 255) / 256), dim3(256), 0, b->stream, d);
    FB_HIP(hipGetLastError());
    FB_HIP(hipStreamSynchronize(b->stream));
    b->inDriven.assign((size_t)N, 0);
    b->drivenList.reserve((size_t)N);   // the driven set never reallocates (fb_apply_deltas_noalloc)
    *out = b;
    return FB_OK;
}

extern "C" fb_status fb_brain_destroy(fb_brain *b) {
    FB_REQUIRE(b, FB_ERR_ARG, "fb_brain_destroy: null brain");
    hipStreamSynchronize(b->stream);
    drop_graphs(b);
    persistent_release(b->d.driven);
    for (void *p : b->allocs) hipFree(p);
    if (b->deltaBuf) hipFree(b->deltaBuf);
    if (b->gatherBuf) hipFree(b->gatherBuf);
    if (b->gatherCnt) hipFree(b->gatherCnt);
    hipStreamDestroy(b->ownStream);
    fb_graph_release(b->g);
    delete b;
    return FB_OK;
}

extern "C" fb_status fb_reset(fb_brain *b) {
    FB_REQUIRE(b, FB_ERR_ARG, "fb_reset: null brain");
    fb_status s0 = sync_stage(b, true, true);
    if (s0) return s0;
    b->gENonneg = true;
    hipLaunchKernelGGL(k_reset_state, dim3((std::max(b->d.N, b->d.nslots) + 255) / 256), dim3(256), 0, b->stream,
                       b->d);
    FB_HIP(hipGetLastError());
    if (b->d.poolOf) {                 // the pool loop starts over with the state
        FB_HIP(hipMemsetAsync(b->d.loopState, 0, sizeof(int32_t) * 2, b->stream));
        FB_HIP(hipMemsetAsync(b->d.loopHist, 0, sizeof(uint32_t) * b->d.loopRing, b->stream));
    }
    b->t = 0.0;
    return FB_OK;
}

extern "C" fb_status fb_set_params(fb_brain *b, const fb_params *p) {
    FB_REQUIRE(b && p, FB_ERR_ARG, "fb_set_params: null argument");
    FB_REQUIRE(p->dt > 0 && p->tauM > 0 && p->tauSyn > 0, FB_ERR_ARG, "fb_set_params: dt, tauM, tauSyn must be > 0");
    fb_status s0 = sync_stage(b, true, true);
    if (s0) return s0;
    FB_HIP(hipStreamSynchronize(b->stream));
    const int32_t ns = nslots_of(*p);
    b->p = *p;
    derive_consts(*p, b->d.N, b->d.k);
    if (ns != b->d.nslots) {
        // as src/lif.js:47 re-creates its delay line when delay or dt change
        b->d.nslots = ns;
        fb_status st = alloc_ring(b);
        if (st) return st;
        b->head = 0;
        b->lastSlot = 0;
    }
    b->version++;
    return FB_OK;
}

extern "C" fb_status fb_get_params(const fb_brain *b, fb_params *p) {
    FB_REQUIRE(b && p, FB_ERR_ARG, "fb_get_params: null argument");
    *p = b->p;
    return FB_OK;
}

static fb_status set_values(fb_brain *b, float *arr, const int32_t *idx, int32_t n, double value, int mode) {
    FB_REQUIRE(n >= 0 && (n == 0 || idx), FB_ERR_ARG, "setter: bad index list");
    for (int32_t q = 0; q < n; q++)
        FB_REQUIRE(idx[q] >= 0 && idx[q] < b->d.N, FB_ERR_ARG, "setter: index %d out of range", idx[q]);
    if (n == 0) return FB_OK;
    fb_status s0 = sync_stage(b, true, true);
    if (s0) return s0;
    if (n == 1) {
        hipLaunchKernelGGL(k_set, dim3(1), dim3(64), 0, b->stream, arr, (const int32_t *)nullptr, 1, idx[0], value,
                           mode, 0);
    } else {
        int serial = 0;
        if (mode == 1) {   // repeated indices must add in list order
            std::vector<int32_t> s(idx, idx + n);
            std::sort(s.begin(), s.end());
            serial = std::adjacent_find(s.begin(), s.end()) != s.end();
        }
        FB_HIP(hipMemcpyAsync(b->idxScratch, idx, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, b->stream));
        hipLaunchKernelGGL(k_set, dim3((n + 255) / 256), dim3(256), 0, b->stream, arr, (const int32_t *)b->idxScratch,
                           n, 0, value, mode, serial);
        FB_HIP(hipStreamSynchronize(b->stream));   // idxScratch is reused by the next setter
    }
    FB_HIP(hipGetLastError());
    return FB_OK;
}

extern "C" fb_status fb_set_drive_one(fb_brain *b, int32_t i, double hz) {
    FB_REQUIRE(b && i >= 0 && i < (b ? b->d.N : 0), FB_ERR_ARG, "fb_set_drive_one: bad neuron");
    fb_status st = set_values(b, b->d.drive, &i, 1, hz, 0);
    if (st) return st;
    // drivenSet semantics (src/lifwasm.js:71): add keeps first-insertion order,
    // removal drops the entry so a later re-add goes to the end
    if (hz > 0) {
        if (!b->inDriven[i]) {
            b->inDriven[i] = 1;
            b->drivenList.push_back(i);
            b->drivenDirty = true;
        }
    } else if (b->inDriven[i]) {
        b->inDriven[i] = 0;
        b->drivenList.erase(std::find(b->drivenList.begin(), b->drivenList.end(), i));
        b->drivenDirty = true;
    }
    return FB_OK;
}

extern "C" fb_status fb_set_drive(fb_brain *b, const int32_t *idx, int32_t n, double hz) {
    FB_REQUIRE(b && n >= 0 && (n == 0 || idx), FB_ERR_ARG, "fb_set_drive: bad arguments");
    for (int32_t q = 0; q < n; q++)
        FB_REQUIRE(idx[q] >= 0 && idx[q] < b->d.N, FB_ERR_ARG, "fb_set_drive: index %d out of range", idx[q]);
    fb_status st = set_values(b, b->d.drive, idx, n, hz, 0);
    if (st) return st;
    for (int32_t q = 0; q < n; q++) {
        const int32_t i = idx[q];
        if (hz > 0) {
            if (!b->inDriven[i]) { b->inDriven[i] = 1; b->drivenList.push_back(i); b->drivenDirty = true; }
        } else if (b->inDriven[i]) {
            b->inDriven[i] = 0;
            b->drivenList.erase(std::find(b->drivenList.begin(), b->drivenList.end(), i));
            b->drivenDirty = true;
        }
    }
    return FB_OK;
}

extern "C" fb_status fb_set_bias(fb_brain *b, const int32_t *idx, int32_t n, double mv) {
    FB_REQUIRE(b, FB_ERR_ARG, "fb_set_bias: null brain");
    return set_values(b, b->d.bias, idx, n, mv, 0);
}

extern "C" fb_status fb_set_thr(fb_brain *b, int32_t i, double mv) {
    FB_REQUIRE(b, FB_ERR_ARG, "fb_set_thr: null brain");
    return set_values(b, b->d.thr, &i, 1, mv, 0);
}

extern "C" fb_status fb_add_g(fb_brain *b, int32_t i, double e, double ii) {
    FB_REQUIRE(b, FB_ERR_ARG, "fb_add_g: null brain");
    if (std::signbit(e)) b->gENonneg = false;
    fb_status st = set_values(b, b->d.gE, &i, 1, e, 1);
    if (st) return st;
    return set_values(b, b->d.gI, &i, 1, ii, 1);
}

extern "C" fb_status fb_pulse(fb_brain *b, const int32_t *idx, int32_t n, double mv) {
    FB_REQUIRE(b, FB_ERR_ARG, "fb_pulse: null brain");
    if (n > 0 && std::signbit(mv)) b->gENonneg = false;
    return set_values(b, b->d.gE, idx, n, mv, 1);
}

// One thread per (array, neuron) group of a delta queue: its operations in
// queue order (sets store f32(value), adds store f32(f64(old) + value)).
__global__ void k_apply_deltas(DevBrain d, const int32_t *gOff, const int32_t *idx, const int32_t *kind,
                               const double *val, int32_t ngroups) {
    const int32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngroups) return;
    const int32_t q0 = gOff[g], q1 = gOff[g + 1];
    const int32_t i = idx[q0];
    float *arr = nullptr;
    switch (kind[q0]) {
        case FB_DELTA_DRIVE: arr = d.drive; break;
        case FB_DELTA_BIAS: arr = d.bias; break;
        case FB_DELTA_THR: arr = d.thr; break;
        case FB_DELTA_GE: arr = d.gE; break;
        default: arr = d.gI; break;
    }
    float x = arr[i];
    for (int32_t q = q0; q < q1; q++) {
        const int32_t k = kind[q];
        x = (k == FB_DELTA_GE || k == FB_DELTA_GI) ? (float)((double)x + val[q]) : (float)val[q];
    }
    arr[i] = x;
}

// the queue's checks and the host side of the setters, in queue order: the
// driven set and the sign bookkeeping of the excitatory conductances (no
// allocation: the driven list has room for every neuron)
static fb_status deltas_host_side(fb_brain *b, int32_t n, const int32_t *idx, const int32_t *kind, const double *val,
                                  const char *who) {
    FB_REQUIRE(b && n >= 0 && (n == 0 || (idx && kind && val)), FB_ERR_ARG, "%s: bad arguments", who);
    for (int32_t q = 0; q < n; q++) {
        FB_REQUIRE(idx[q] >= 0 && idx[q] < b->d.N, FB_ERR_ARG, "%s: index %d out of range", who, idx[q]);
        FB_REQUIRE(kind[q] >= FB_DELTA_DRIVE && kind[q] <= FB_DELTA_GI, FB_ERR_ARG, "%s: bad kind %d", who, kind[q]);
    }
    for (int32_t q = 0; q < n; q++) {
        const int32_t i = idx[q];
        if (kind[q] == FB_DELTA_DRIVE) {
            if (val[q] > 0) {
                if (!b->inDriven[i]) { b->inDriven[i] = 1; b->drivenList.push_back(i); b->drivenDirty = true; }
            } else if (b->inDriven[i]) {
                b->inDriven[i] = 0;
                b->drivenList.erase(std::find(b->drivenList.begin(), b->drivenList.end(), i));
                b->drivenDirty = true;
            }
        } else if (kind[q] == FB_DELTA_GE && std::signbit(val[q])) {
            b->gENonneg = false;
        }
    }
    return FB_OK;
}

static int32_t delta_array_of(int32_t k) { return k <= FB_DELTA_THR ? k : (k == FB_DELTA_GE ? 3 : 4); }

extern "C" fb_status fb_apply_deltas(fb_brain *b, int32_t n, const int32_t *idx, const int32_t *kind,
                                     const double *val) {
    fb_status st0 = deltas_host_side(b, n, idx, kind, val, "fb_apply_deltas");
    if (st0 || n == 0) return st0;
    if ((st0 = sync_stage(b, true, true))) return st0;
    // group by (array, neuron), keeping queue order inside a group
    auto arrayOf = delta_array_of;
    std::vector<int32_t> perm((size_t)n);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int32_t x, int32_t y) {
        const int32_t ax = arrayOf(kind[x]), ay = arrayOf(kind[y]);
        return ax != ay ? ax < ay : idx[x] < idx[y];
    });
    std::vector<int32_t> sIdx((size_t)n), sKind((size_t)n), gOff;
    std::vector<double> sVal((size_t)n);
    for (int32_t q = 0; q < n; q++) {
        const int32_t p = perm[(size_t)q];
        sIdx[(size_t)q] = idx[p]; sKind[(size_t)q] = kind[p]; sVal[(size_t)q] = val[p];
        if (q == 0 || idx[p] != sIdx[(size_t)q - 1] || arrayOf(kind[p]) != arrayOf(sKind[(size_t)q - 1]))
            gOff.push_back(q);
    }
    const int32_t ng = (int32_t)gOff.size();
    gOff.push_back(n);
    if ((int64_t)n > b->deltaCap) {
        if (b->deltaBuf) FB_HIP(hipFree(b->deltaBuf));
        b->deltaBuf = nullptr;
        b->deltaCap = std::max<int64_t>((int64_t)n, 4096);
        FB_HIP(hipMalloc(&b->deltaBuf, (size_t)b->deltaCap * (sizeof(double) + 3 * sizeof(int32_t)) + 4 * sizeof(int32_t)));
    }
    double *dVal = (double *)b->deltaBuf;
    int32_t *dIdx = (int32_t *)(dVal + b->deltaCap), *dKind = dIdx + b->deltaCap, *dOff = dKind + b->deltaCap;
    FB_HIP(hipMemcpyAsync(dVal, sVal.data(), sizeof(double) * (size_t)n, hipMemcpyHostToDevice, b->stream));
    FB_HIP(hipMemcpyAsync(dIdx, sIdx.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, b->stream));
    FB_HIP(hipMemcpyAsync(dKind, sKind.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, b->stream));
    FB_HIP(hipMemcpyAsync(dOff, gOff.data(), sizeof(int32_t) * ((size_t)ng + 1), hipMemcpyHostToDevice, b->stream));
    hipLaunchKernelGGL(k_apply_deltas, dim3((ng + 255) / 256), dim3(256), 0, b->stream, b->d, (const int32_t *)dOff,
                       (const int32_t *)dIdx, (const int32_t *)dKind, (const double *)dVal, ng);
    FB_HIP(hipGetLastError());
    FB_HIP(hipStreamSynchronize(b->stream));   // the host vectors and the staging buffer are reused
    return FB_OK;
}

// ------------------------------------------------------------------ allocation-free deltas
//
// fb_apply_deltas_noalloc: the same queue applied with no device or host
// allocation, for callers that cannot allocate (code under HIP graph capture,
// tight serving loops).  The queue's staging lives in `scratch`, one array
// per device for the life of the process; the grouping sort in host arrays of
// the same bound.  Every call on a device shares them, so calls that overlap on
// one GPU race (a mutex orders calls from this process's threads).  No part of
// the staging needs clearing: a call writes every word it reads.
constexpr int32_t NOALLOC_MAX_DELTAS = 65536;   // LIFGpu's delta queue bound (src/lifgpu.js)
constexpr size_t NOALLOC_VAL = 0, NOALLOC_IDX = NOALLOC_VAL + sizeof(double) * NOALLOC_MAX_DELTAS,
                 NOALLOC_KIND = NOALLOC_IDX + sizeof(int32_t) * NOALLOC_MAX_DELTAS,
                 NOALLOC_OFF = NOALLOC_KIND + sizeof(int32_t) * NOALLOC_MAX_DELTAS,
                 NOALLOC_BYTES = NOALLOC_OFF + sizeof(int32_t) * (NOALLOC_MAX_DELTAS + 1);
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_noallocMu;
static unsigned char *g_scratchPtr[NOALLOC_MAX_DEVICES];   // scratch's device address, per device
static uint64_t g_key[NOALLOC_MAX_DELTAS];
static int32_t g_sIdx[NOALLOC_MAX_DELTAS], g_sKind[NOALLOC_MAX_DELTAS], g_gOff[NOALLOC_MAX_DELTAS + 1];
static double g_sVal[NOALLOC_MAX_DELTAS];

extern "C" fb_status fb_apply_deltas_noalloc(fb_brain *b, int32_t n, const int32_t *idx, const int32_t *kind,
                                             const double *val) {
    FB_REQUIRE(n <= NOALLOC_MAX_DELTAS, FB_ERR_ARG, "fb_apply_deltas_noalloc: %d deltas exceed the maximum of %d", n,
               NOALLOC_MAX_DELTAS);
    std::lock_guard<std::mutex> lock(g_noallocMu);
    fb_status st0 = deltas_host_side(b, n, idx, kind, val, "fb_apply_deltas_noalloc");
    if (st0 || n == 0) return st0;
    if ((st0 = sync_stage(b, true, true))) return st0;   // a kernel launch at most, no allocation
    const int dev = b->g->device;
    FB_REQUIRE(dev >= 0 && dev < NOALLOC_MAX_DEVICES, FB_ERR_UNSUPPORTED, "fb_apply_deltas_noalloc: device %d", dev);
    if (!g_scratchPtr[dev]) {
        // once per device, with that device current: the array's own address
        // on it (the name itself is only a host-side stand-in)
        int cur = 0;
        FB_HIP(hipGetDevice(&cur));
        if (cur != dev) FB_HIP(hipSetDevice(dev));
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != dev) FB_HIP(hipSetDevice(cur));
        FB_REQUIRE(e == hipSuccess && ptr, FB_ERR_DEVICE, "fb_apply_deltas_noalloc: hipGetSymbolAddress failed: %s",
                   hipGetErrorString(e));
        g_scratchPtr[dev] = (unsigned char *)ptr;
    }
    // group by (array, neuron) keeping queue order: one in-place sort of
    // array (3 bits) | neuron (31 bits) | queue position (17 bits)
    for (int32_t q = 0; q < n; q++)
        g_key[q] = ((uint64_t)delta_array_of(kind[q]) << 48) | ((uint64_t)(uint32_t)idx[q] << 17) | (uint64_t)q;
    std::sort(g_key, g_key + n);
    int32_t ng = 0;
    for (int32_t q = 0; q < n; q++) {
        const int32_t p = (int32_t)(g_key[q] & 0x1FFFF);
        g_sIdx[q] = idx[p]; g_sKind[q] = kind[p]; g_sVal[q] = val[p];
        if (q == 0 || (g_key[q] >> 17) != (g_key[q - 1] >> 17)) g_gOff[ng++] = q;
    }
    g_gOff[ng] = n;
    unsigned char *base = g_scratchPtr[dev];
    double *dVal = (double *)(base + NOALLOC_VAL);
    int32_t *dIdx = (int32_t *)(base + NOALLOC_IDX), *dKind = (int32_t *)(base + NOALLOC_KIND),
            *dOff = (int32_t *)(base + NOALLOC_OFF);
    FB_HIP(hipMemcpyAsync(dVal, g_sVal, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, b->stream));
    FB_HIP(hipMemcpyAsync(dIdx, g_sIdx, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, b->stream));
    FB_HIP(hipMemcpyAsync(dKind, g_sKind, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, b->stream));
    FB_HIP(hipMemcpyAsync(dOff, g_gOff, sizeof(int32_t) * ((size_t)ng + 1), hipMemcpyHostToDevice, b->stream));
    hipLaunchKernelGGL(k_apply_deltas, dim3((ng + 255) / 256), dim3(256), 0, b->stream, b->d, (const int32_t *)dOff,
                       (const int32_t *)dIdx, (const int32_t *)dKind, (const double *)dVal, ng);
    FB_HIP(hipGetLastError());
    FB_HIP(hipStreamSynchronize(b->stream));   // the host arrays and the staging are shared by the next call
    return FB_OK;
}

extern "C" fb_status fb_set_background(fb_brain *b, double hz, double mv) {
    FB_REQUIRE(b, FB_ERR_ARG, "fb_set_background: null brain");
    fb_status s0 = sync_stage(b, true, true);
    if (s0) return s0;
    b->p.bgRate = hz;
    b->p.bgAmp = mv;
    b->d.k.bgEvents = (float)((double)b->d.N * hz * b->p.dt / 1000.0);
    b->d.k.bgAmp = (float)mv;
    b->version++;
    return FB_OK;
}

extern "C" fb_status fb_driven(const fb_brain *b, int32_t *out, int32_t cap, int32_t *n) {
    FB_REQUIRE(b && n, FB_ERR_ARG, "fb_driven: null argument");
    *n = (int32_t)b->drivenList.size();
    if (out) for (int32_t q = 0; q < std::min(cap, *n); q++) out[q] = b->drivenList[q];
    return FB_OK;
}

// ------------------------------------------------------------------ stepping

// lif.c's deliver loop is order-dependent beyond what the parallel scatter
// reproduces in two cases: presynaptic inhibition reads gI of the arriving
// neuron mid-delivery, and with tRef + dt - dt <= 0 a forced neuron can also
// cross threshold in its own step, so one neuron arrives twice in one list
// and its depression is applied twice in order.  Both take the serial path.
static bool serial_deliver(const fb_brain *b) {
    const Consts &k = b->d.k;
    const float forced = k.tRef + k.dt;
    return k.preInh > 0.f || !(forced - k.dt > 0.f) || !(k.tRef > 0.f);
}

static fb_status check_device_errors(fb_brain *b) {
    Scalars sc;
    FB_HIP(hipMemcpyAsync(&sc, b->d.sc, sizeof sc, hipMemcpyDeviceToHost, b->stream));
    FB_HIP(hipStreamSynchronize(b->stream));
    FB_REQUIRE(!(sc.err & FB_DEVERR_RUN_OVERFLOW), FB_ERR_OVERFLOW, "a target received more arrivals in one step than it has in-edges");
    return FB_OK;
}

static fb_status run_launch(fb_brain *b, int32_t nsteps) {
    for (int32_t s = 0; s < nsteps; s++) {
        const int32_t slot = (b->head + b->d.nslots - 1) % b->d.nslots;
        fb_status st = launch_step(b->d, b->head, slot, serial_deliver(b), b->g->rowsUnique, b->stream);
        if (st) return st;
        b->lastSlot = slot;
        b->head = (b->head + 1) % b->d.nslots;
    }
    return FB_OK;
}

// A hipGraph holds a whole number of slot cycles starting at a given head;
// after a replay the head is back where it started, so the same graph
// replays again.
static fb_status run_graph(fb_brain *b, int32_t nsteps) {
    const int32_t ns = b->d.nslots;
    if (b->graphVersion != b->version) {
        drop_graphs(b);
        b->graphVersion = b->version;
    }
    // long runs replay graphs of at least 64 steps (whole slot cycles), so the
    // host's launch gap between replays is paid rarely
    const int32_t big = ns * std::max(1, (64 + ns - 1) / ns);
    while (nsteps >= ns) {
        const int32_t len = nsteps >= big ? big : ns;
        GraphCache &gc = b->graphs[{b->head, len}];
        if (!gc.exec) {
            FB_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            int32_t h = b->head;
            for (int32_t s = 0; s < len; s++) {
                const int32_t slot = (h + ns - 1) % ns;
                fb_status st = launch_step(b->d, h, slot, serial_deliver(b), b->g->rowsUnique, b->stream);
                if (st) { hipGraph_t tmp; hipStreamEndCapture(b->stream, &tmp); if (tmp) hipGraphDestroy(tmp); return st; }
                h = (h + 1) % ns;
            }
            FB_HIP(hipStreamEndCapture(b->stream, &gc.graph));
            FB_HIP(hipGraphInstantiate(&gc.exec, gc.graph, nullptr, nullptr, 0));
            gc.head = b->head;
            gc.steps = len;
        }
        FB_HIP(hipGraphLaunch(gc.exec, b->stream));
        b->lastSlot = (b->head + len - 1 + ns - 1) % ns;   // out slot of the last step in the replay
        nsteps -= len;
    }
    return run_launch(b, nsteps);
}

static fb_status run_steps(fb_brain *b, int32_t nsteps, int32_t engine) {
    FB_REQUIRE(nsteps >= 0, FB_ERR_ARG, "fb_step: nsteps must be >= 0");
    fb_status st = sync_driven(b);
    if (st) return st;
    if (nsteps == 0) return FB_OK;
    // background events of negative amplitude can leave an excitatory conductance at -0
    if (b->d.k.bgEvents > 0.f && std::signbit(b->d.k.bgAmp)) b->gENonneg = false;
    const int32_t skip = b->g->wNonneg && b->gENonneg ? 1 : 0;
    if (skip != b->d.skipZeroW) {
        b->d.skipZeroW = skip;
        drop_graphs(b);   // captured launches hold the old value
    }
    const bool autoChosen = engine == FB_ENGINE_AUTO;
    if (engine == FB_ENGINE_AUTO) {
        // the lookahead engine pays a prologue per call (one launch that
        // loads the state and scatters the delay line).  Measured on MI355X
        // (events around a call, in a loop of such calls): 57 us + 2.7 us a
        // step at 165,122 neurons against 19 us a step per-step, 44 us + 1.5
        // us a step at 16,384 against 12 us; so from 4 steps on for brains of
        // that size, from one epoch of steps on for smaller ones (whose
        // per-step kernels are shorter)
        std::string why;
        const bool longRun = nsteps >= b->d.nslots - 1 || (nsteps >= 4 && b->d.N >= 16384);
        // otherwise per-step kernels: replayed as a graph for small brains,
        // whose kernels are shorter than their launch; launched directly for
        // large ones, whose kernels hide the launch and where a capture after
        // every drive change would cost more than it saves
        engine = longRun && persistent_supported(b->d, serial_deliver(b), &why) == FB_OK ? FB_ENGINE_PERSISTENT
                 : b->d.N >= 65536                                                       ? FB_ENGINE_LAUNCH
                                                                                          : FB_ENGINE_GRAPH;
    }
    if (engine != FB_ENGINE_PERSISTENT) {
        // the per-step engines step the brain's arrays
        st = sync_stage(b, true, false);
        if (st) return st;
    }
    switch (engine) {
        case FB_ENGINE_LAUNCH: st = run_launch(b, nsteps); break;
        case FB_ENGINE_GRAPH: st = run_graph(b, nsteps); break;
        case FB_ENGINE_PERSISTENT: {
            std::string why;
            st = persistent_supported(b->d, serial_deliver(b), &why);
            if (st) { set_error("persistent engine: %s", why.c_str()); return st; }
            st = launch_persistent(b->d, b->head, nsteps, b->stream, &b->fallbackSteps, &b->persistentLaunches,
                                   &b->persistentMs, b->phaseUs, b->drivenVersion);
            if (st == FB_ERR_UNSUPPORTED && autoChosen) {
                // the device refused the launch before any step ran
                st = sync_stage(b, true, false);
                if (st) return st;
                st = b->d.N >= 65536 ? run_launch(b, nsteps) : run_graph(b, nsteps);
                break;
            }
            if (st) return st;
            b->lastSlot = (b->head + nsteps - 1 + b->d.nslots - 1) % b->d.nslots;
            b->head = (b->head + nsteps) % b->d.nslots;
            break;
        }
        default: set_error("fb_step: unknown engine %d", engine); return FB_ERR_ARG;
    }
    if (st) return st;
    b->t += (double)nsteps * b->p.dt;
    b->steps += nsteps;
    return FB_OK;
}

extern "C" fb_status fb_step(fb_brain *b, int32_t nsteps, int32_t engine, int32_t *nfired_last) {
    FB_REQUIRE(b, FB_ERR_ARG, "fb_step: null brain");
    fb_status st = run_steps(b, nsteps, engine);
    if (st) return st;
    if (nfired_last) {
        FB_HIP(hipMemcpyAsync(nfired_last, b->d.ringCount + b->lastSlot, sizeof(int32_t), hipMemcpyDeviceToHost,
                              b->stream));
        FB_HIP(hipStreamSynchronize(b->stream));
        return check_device_errors(b);
    }
    return FB_OK;
}

extern "C" fb_status fb_brain_stat(fb_brain *b, const char *name, double *value) {
    FB_REQUIRE(b && name && value, FB_ERR_ARG, "fb_brain_stat: null argument");
    std::string n(name);
    std::string why;
    if (n == "fallback_steps") *value = (double)b->fallbackSteps;
    else if (n == "persistent_launches") *value = (double)b->persistentLaunches;
    else if (n == "persistent_kernel_ms") {
        persistent_take_time(b->d.driven, &b->persistentMs);
        *value = b->persistentMs;
    }
    else if (n == "persistent_overflow_flags") *value = (double)persistent_overflow_flags(b->d);
    else if (n.rfind("phase_us_", 0) == 0) {
        // phase_us_<drv|reg|spare>_<phase>: time the driven block, the first
        // regular block and the first spare block spent in each phase of the
        // lookahead kernel
        static const char *ph[NPH] = {"gather", "sim", "push", "lists", "barrier", "prologue", "share", "pushwait"};
        static const char *blkName[3] = {"drv_", "reg_", "spare_"};
        int idx = -1;
        for (int blk = 0; blk < 3; blk++)
            for (int k = 0; k < NPH; k++)
                if (n == std::string("phase_us_") + blkName[blk] + ph[k]) idx = blk * NPH + k;
        static const char *sub[6] = {"inputs", "scan", "draws", "decide", "membrane", "sync"};
        for (int k = 0; k < 6; k++)
            if (n == std::string("phase_us_drvstep_") + sub[k]) idx = PROF_SUB + k;
        // the shared push of the driven rows: mean per block, and the slowest block
        static const char *sh[4] = {"load", "edges", "publish", "max"};
        for (int k = 0; k < 4; k++)
            if (n == std::string("phase_us_share_") + sh[k]) idx = PROF_SHARE + k;
        // the launch's span: spread of the block starts, the first start to the
        // last block's end, and the last block out's hand-over to the host
        // (and the first start to the last end of a driven, regular, spare block)
        static const char *sp[7] = {"start_spread", "body", "handover", "", "end_drv", "end_reg", "end_spare"};
        for (int k = 0; k < 7; k++)
            if (k != 3 && n == std::string("phase_us_span_") + sp[k]) idx = PROF_SPAN + k;
        if (idx < 0) { set_error("fb_brain_stat: unknown statistic '%s'", name); return FB_ERR_ARG; }
        *value = b->phaseUs[idx];
    }
    else if (n == "persistent_supported")
        *value = persistent_supported(b->d, serial_deliver(b), &why) == FB_OK ? 1.0 : 0.0;
    else if (n == "nslots") *value = (double)b->d.nslots;
    else if (n == "serial_deliver") *value = serial_deliver(b) ? 1.0 : 0.0;
    else { set_error("fb_brain_stat: unknown statistic '%s'", name); return FB_ERR_ARG; }
    return FB_OK;
}

extern "C" fb_status fb_brain_set_stream(fb_brain *b, void *stream) {
    FB_REQUIRE(b, FB_ERR_ARG, "fb_brain_set_stream: null brain");
    FB_HIP(hipStreamSynchronize(b->stream));
    b->stream = stream ? (hipStream_t)stream : b->ownStream;
    drop_graphs(b);   // graphs were captured on the previous stream's work queue
    return FB_OK;
}

extern "C" fb_status fb_synchronize(fb_brain *b) {
    FB_REQUIRE(b, FB_ERR_ARG, "fb_synchronize: null brain");
    FB_HIP(hipStreamSynchronize(b->stream));
    return check_device_errors(b);
}

extern "C" fb_status fb_fired(fb_brain *b, int32_t *out, int32_t cap, int32_t *n) {
    FB_REQUIRE(b && n, FB_ERR_ARG, "fb_fired: null argument");
    int32_t cnt = 0;
    FB_HIP(hipMemcpyAsync(&cnt, b->d.ringCount + b->lastSlot, sizeof(int32_t), hipMemcpyDeviceToHost, b->stream));
    FB_HIP(hipStreamSynchronize(b->stream));
    *n = cnt;
    if (out && cap > 0 && cnt > 0) {
        FB_HIP(hipMemcpyAsync(out, b->d.ring + (int64_t)b->lastSlot * b->d.ringStride, sizeof(int32_t) * (size_t)std::min(cap, cnt),
                              hipMemcpyDeviceToHost, b->stream));
        FB_HIP(hipStreamSynchronize(b->stream));
    }
    return FB_OK;
}

extern "C" fb_status fb_run_record(fb_brain *b, int32_t nsteps, int32_t engine, int32_t *nfired, int32_t *raster,
                                   int64_t cap, int64_t *total) {
    FB_REQUIRE(b && nsteps >= 0 && nfired && total, FB_ERR_ARG, "fb_run_record: bad arguments");
    const int32_t ns = b->d.nslots, N = b->d.N;
    const int32_t chunk = std::max(1, ns - 1);   // a list is delivered (and its count cleared) ns-1 steps later
    if (!b->gatherBuf) {
        FB_HIP(hipMalloc(&b->gatherBuf, sizeof(int32_t) * (size_t)chunk * 2 * (size_t)N));
        FB_HIP(hipMalloc(&b->gatherCnt, sizeof(int32_t) * (size_t)ns));
    }
    int64_t tot = 0;
    std::vector<int32_t> cnts((size_t)chunk);
    for (int32_t done = 0; done < nsteps;) {
        const int32_t m = std::min(chunk, nsteps - done);
        const int32_t firstSlot = (b->head + ns - 1) % ns;
        fb_status st = run_steps(b, m, engine);
        if (st) return st;
        st = launch_gather(b->d, firstSlot, m, b->gatherCnt, b->gatherBuf, b->stream);
        if (st) return st;
        FB_HIP(hipMemcpyAsync(cnts.data(), b->gatherCnt, sizeof(int32_t) * (size_t)m, hipMemcpyDeviceToHost, b->stream));
        FB_HIP(hipStreamSynchronize(b->stream));
        int64_t sum = 0;
        for (int32_t q = 0; q < m; q++) { nfired[done + q] = cnts[q]; sum += cnts[q]; }
        if (raster && tot < cap && sum > 0) {
            const int64_t take = std::min<int64_t>(sum, cap - tot);
            FB_HIP(hipMemcpyAsync(raster + tot, b->gatherBuf, sizeof(int32_t) * (size_t)take, hipMemcpyDeviceToHost,
                                  b->stream));
            FB_HIP(hipStreamSynchronize(b->stream));
        }
        tot += sum;
        done += m;
    }
    *total = tot;
    return check_device_errors(b);
}

extern "C" fb_status fb_get_array(fb_brain *b, int32_t which, void *dst) {
    FB_REQUIRE(b && dst, FB_ERR_ARG, "fb_get_array: null argument");
    float *src = array_ptr(b, which);
    FB_REQUIRE(src, FB_ERR_ARG, "fb_get_array: unknown array %d", which);
    fb_status s0 = sync_stage(b, false, false);
    if (s0) return s0;
    FB_HIP(hipMemcpyAsync(dst, src, sizeof(float) * (size_t)b->d.N, hipMemcpyDeviceToHost, b->stream));
    FB_HIP(hipStreamSynchronize(b->stream));
    return FB_OK;
}

extern "C" fb_status fb_set_array(fb_brain *b, int32_t which, const void *src) {
    FB_REQUIRE(b && src, FB_ERR_ARG, "fb_set_array: null argument");
    float *dst = array_ptr(b, which);
    FB_REQUIRE(dst, FB_ERR_ARG, "fb_set_array: unknown array %d", which);
    fb_status s0 = sync_stage(b, true, true);
    if (s0) return s0;
    if (dst == b->d.gE) {
        const float *v = (const float *)src;
        bool nonneg = true;
        for (int32_t i = 0; i < b->d.N; i++) nonneg = nonneg && !std::signbit(v[i]);
        b->gENonneg = nonneg;
    }
    FB_HIP(hipMemcpyAsync(dst, src, sizeof(float) * (size_t)b->d.N, hipMemcpyHostToDevice, b->stream));
    FB_HIP(hipStreamSynchronize(b->stream));
    return FB_OK;
}

extern "C" fb_status fb_set_rng(fb_brain *b, uint32_t rng) {
    FB_REQUIRE(b && rng != 0, FB_ERR_ARG, "fb_set_rng: null brain or a zero state");
    FB_HIP(hipMemcpyAsync(&b->d.sc->rng, &rng, sizeof rng, hipMemcpyHostToDevice, b->stream));
    FB_HIP(hipStreamSynchronize(b->stream));
    return FB_OK;
}

// frees one of the brain's own device buffers before its end of life
static void free_owned(fb_brain *b, const void *p) {
    if (!p) return;
    auto it = std::find(b->allocs.begin(), b->allocs.end(), (void *)p);
    if (it == b->allocs.end()) return;
    b->allocs.erase(it);
    (void)hipFree((void *)p);
}

extern "C" fb_status fb_set_pool_feedback(fb_brain *b, const uint8_t *poolOf, int32_t nPools, const int32_t *targets,
                                          int32_t nTargets, float base, float gain, int32_t delay,
                                          int32_t recordCap) {
    FB_REQUIRE(b, FB_ERR_ARG, "fb_set_pool_feedback: null brain");
    fb_status s0 = sync_stage(b, true, true);
    if (s0) return s0;
    FB_HIP(hipStreamSynchronize(b->stream));
    if (b->d.poolOf) {                 // the previous loop's buffers
        DevBrain &o = b->d;
        free_owned(b, o.poolOf); free_owned(b, o.loopTargets); free_owned(b, o.loopState);
        free_owned(b, o.loopHist); free_owned(b, o.loopRec);
        o.poolOf = nullptr; o.loopTargets = nullptr; o.loopState = nullptr; o.loopHist = nullptr; o.loopRec = nullptr;
        drop_graphs(b);                // captured launches hold the freed pointers
        b->version++;
    }
    if (!poolOf) return FB_OK;
    FB_REQUIRE(nPools >= 1 && nPools <= 8 && nTargets >= 0 && (nTargets == 0 || targets) && delay >= 1 &&
                   recordCap >= 0,
               FB_ERR_ARG, "fb_set_pool_feedback: bad arguments");
    for (int32_t q = 0; q < nTargets; q++)
        FB_REQUIRE(targets[q] >= 0 && targets[q] < b->d.N, FB_ERR_ARG, "fb_set_pool_feedback: target %d out of range",
                   q);
    const int32_t N = b->d.N, ring = delay + 1;
    uint8_t *pool;
    int32_t *tg, *state;
    uint32_t *hist;
    uint16_t *rec;
    fb_status st;
    if ((st = dalloc(b->allocs, &pool, (size_t)N)) || (st = dalloc(b->allocs, &tg, (size_t)std::max(nTargets, 1))) ||
        (st = dalloc(b->allocs, &state, 2)) || (st = dalloc(b->allocs, &hist, (size_t)ring)) ||
        (st = dalloc(b->allocs, &rec, (size_t)std::max(recordCap, 1) * nPools)))
        return st;
    FB_HIP(hipMemcpy(pool, poolOf, (size_t)N, hipMemcpyHostToDevice));
    if (nTargets) FB_HIP(hipMemcpy(tg, targets, sizeof(int32_t) * nTargets, hipMemcpyHostToDevice));
    FB_HIP(hipMemset(state, 0, sizeof(int32_t) * 2));
    FB_HIP(hipMemset(hist, 0, sizeof(uint32_t) * ring));
    DevBrain &d = b->d;
    d.poolOf = pool; d.loopTargets = tg; d.loopHist = hist; d.loopRec = rec; d.loopState = state;
    d.loopPools = nPools; d.loopNTargets = nTargets; d.loopDelay = delay; d.loopRing = ring;
    d.loopRecCap = recordCap; d.loopBase = base; d.loopGain = gain;
    b->version++;                      // captured launches hold the old arguments
    return FB_OK;
}

extern "C" fb_status fb_pool_counts(fb_brain *b, uint16_t *dst, int32_t cap, int32_t *nRecords) {
    FB_REQUIRE(b && nRecords && (cap == 0 || dst), FB_ERR_ARG, "fb_pool_counts: bad arguments");
    FB_REQUIRE(b->d.poolOf, FB_ERR_ARG, "fb_pool_counts: no pool loop is set");
    int32_t state[2];
    FB_HIP(hipMemcpyAsync(state, b->d.loopState, sizeof state, hipMemcpyDeviceToHost, b->stream));
    FB_HIP(hipStreamSynchronize(b->stream));
    *nRecords = state[1];
    const int32_t n = std::min(cap, state[1]);
    if (n > 0)
        FB_HIP(hipMemcpy(dst, b->d.loopRec, sizeof(uint16_t) * (size_t)n * b->d.loopPools, hipMemcpyDeviceToHost));
    return FB_OK;
}

extern "C" fb_status fb_brain_scalars(fb_brain *b, uint32_t *rng, int32_t *head, float *bg_acc, double *t_ms) {
    FB_REQUIRE(b, FB_ERR_ARG, "fb_brain_scalars: null brain");
    Scalars sc;
    FB_HIP(hipMemcpyAsync(&sc, b->d.sc, sizeof sc, hipMemcpyDeviceToHost, b->stream));
    FB_HIP(hipStreamSynchronize(b->stream));
    if (rng) *rng = sc.rng;
    if (head) *head = b->head;
    if (bg_acc) *bg_acc = sc.bgAcc;
    if (t_ms) *t_ms = b->t;
    return FB_OK;
}

extern "C" fb_status fb_brain_size(const fb_brain *b, int32_t *N) {
    FB_REQUIRE(b && N, FB_ERR_ARG, "fb_brain_size: bad arguments");
    *N = b->d.N;
    return FB_OK;
}

extern "C" fb_status fb_brain_counters(fb_brain *b, int64_t *edge_events, int64_t *synapse_events, int64_t *spikes,
                                       int64_t *steps) {
    FB_REQUIRE(b, FB_ERR_ARG, "fb_brain_counters: null brain");
    Scalars sc;
    FB_HIP(hipMemcpyAsync(&sc, b->d.sc, sizeof sc, hipMemcpyDeviceToHost, b->stream));
    FB_HIP(hipStreamSynchronize(b->stream));
    if (edge_events) *edge_events = (int64_t)sc.edgeEvents;
    if (synapse_events) *synapse_events = (int64_t)sc.synEvents;
    if (spikes) *spikes = (int64_t)sc.spikes;
    if (steps) *steps = b->steps;
    return FB_OK;
}
