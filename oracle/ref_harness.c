/*
 * ref_harness.c -- drives the reference's own lif_step, compiled natively
 * from /root/reference/src/wasm/lif.c by oracle/Makefile into oracle/_ref/.
 *
 * TEST INFRASTRUCTURE ONLY.  It exists to pin the restatement in
 * lif_oracle.c: the golden vectors under tests/golden/ are what this harness
 * produced, and the restatement must reproduce them bit for bit.
 *
 * The struct below is the reference's brain header layout (src/wasm/lif.c:4-25),
 * declared here because lif.c defines it privately; the field order and types
 * must stay identical or lif_step reads garbage.  Constants, seed hash and
 * nslots come from the JS wrapper's rules (src/lifwasm.js:39,52,55-68), which
 * lif_oracle.c restates and this file reuses.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "lif_oracle.h"

typedef struct {
    int32_t N, nslots, head, coba;
    float dt, vRest, vThresh, vReset, tRef, adaptInc, depU;
    float dE, dTr, dA, kRec, kM, dtS, eExc, eInh, cE, cI;
    uint32_t rng;
    float *v, *gE, *gI, *refr, *trace, *adapt, *res, *bias, *thr, *drive, *sign;
    uint32_t *spikeCount;
    const uint32_t *indptr, *indices;
    const float *weights;
    int32_t *ring, *ringCount;
    int32_t *driven;
    int32_t nDriven;
    int32_t nFired;
    int32_t lastSlot;
    float bgEvents, bgAmp, bgAcc;
    float preInh;
} ref_header;

int32_t lif_step(ref_header *b); /* from the reference source */

typedef struct {
    ref_header h;
    float *arrays[FBO_NARRAYS];
    uint32_t *ip, *ix;
    float *w, *sg;
} ref_brain;

static void ref_fill(ref_brain *r) {
    ref_header *h = &r->h;
    for (int32_t i = 0; i < h->N; i++) {
        h->v[i] = h->vRest; h->gE[i] = 0.f; h->gI[i] = 0.f; h->refr[i] = 0.f;
        h->trace[i] = 0.f; h->adapt[i] = 0.f; h->res[i] = 1.f; h->spikeCount[i] = 0;
    }
    memset(h->ringCount, 0, sizeof(int32_t) * (size_t)h->nslots);
}

ref_brain *ref_create(int32_t N, int64_t E, const uint32_t *indptr, const uint32_t *indices,
                      const float *W, const float *sign, const fbo_params *p, double seed) {
    ref_brain *r = (ref_brain *)calloc(1, sizeof(ref_brain));
    ref_header *h = &r->h;
    fbo_consts c;
    fbo_constants(p, N, &c);
    h->N = N; h->nslots = fbo_nslots(p->delay, p->dt); h->head = 0; h->coba = c.coba;
    h->dt = c.dt; h->vRest = c.vRest; h->vThresh = c.vThresh; h->vReset = c.vReset;
    h->tRef = c.tRef; h->adaptInc = c.adaptInc; h->depU = c.depU;
    h->dE = c.dE; h->dTr = c.dTr; h->dA = c.dA; h->kRec = c.kRec; h->kM = c.kM; h->dtS = c.dtS;
    h->eExc = c.eExc; h->eInh = c.eInh; h->cE = c.cE; h->cI = c.cI;
    h->rng = fbo_seed_hash(seed);
    for (int a = 0; a < FBO_NARRAYS; a++) r->arrays[a] = (float *)calloc((size_t)N, sizeof(float));
    h->v = r->arrays[FBO_V]; h->gE = r->arrays[FBO_GE]; h->gI = r->arrays[FBO_GI];
    h->refr = r->arrays[FBO_REFR]; h->trace = r->arrays[FBO_TRACE]; h->adapt = r->arrays[FBO_ADAPT];
    h->res = r->arrays[FBO_RES]; h->bias = r->arrays[FBO_BIAS]; h->thr = r->arrays[FBO_THR];
    h->drive = r->arrays[FBO_DRIVE];
    r->ip = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)N + 1));
    r->ix = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(E ? E : 1));
    r->w = (float *)malloc(sizeof(float) * (size_t)(E ? E : 1));
    r->sg = (float *)malloc(sizeof(float) * (size_t)N);
    memcpy(r->ip, indptr, sizeof(uint32_t) * ((size_t)N + 1));
    if (E) { memcpy(r->ix, indices, sizeof(uint32_t) * (size_t)E); memcpy(r->w, W, sizeof(float) * (size_t)E); }
    memcpy(r->sg, sign, sizeof(float) * (size_t)N);
    h->sign = r->sg; h->indptr = r->ip; h->indices = r->ix; h->weights = r->w;
    h->spikeCount = (uint32_t *)calloc((size_t)N, sizeof(uint32_t));
    h->ring = (int32_t *)calloc((size_t)h->nslots * (size_t)N, sizeof(int32_t));
    h->ringCount = (int32_t *)calloc((size_t)h->nslots, sizeof(int32_t));
    h->driven = (int32_t *)calloc((size_t)N, sizeof(int32_t));
    h->nDriven = 0; h->nFired = 0; h->lastSlot = 0;
    h->bgEvents = c.bgEvents; h->bgAmp = c.bgAmp; h->bgAcc = 0.f; h->preInh = c.preInh;
    ref_fill(r);
    return r;
}

void ref_destroy(ref_brain *r) {
    if (!r) return;
    for (int a = 0; a < FBO_NARRAYS; a++) free(r->arrays[a]);
    free(r->ip); free(r->ix); free(r->w); free(r->sg);
    free(r->h.spikeCount); free(r->h.ring); free(r->h.ringCount); free(r->h.driven);
    free(r);
}

void ref_reset(ref_brain *r) { ref_fill(r); }
void ref_set_driven(ref_brain *r, const int32_t *list, int32_t n) {
    for (int32_t k = 0; k < n; k++) r->h.driven[k] = list[k];
    r->h.nDriven = n;
}
float *ref_array(ref_brain *r, int32_t which) { return r->arrays[which]; }
uint32_t *ref_spike_count(ref_brain *r) { return r->h.spikeCount; }
uint32_t ref_rng_state(const ref_brain *r) { return r->h.rng; }
void ref_set_rng(ref_brain *r, uint32_t rng) { r->h.rng = rng; }
int32_t ref_head(const ref_brain *r) { return r->h.head; }
int32_t ref_nslots_of(const ref_brain *r) { return r->h.nslots; }
float ref_bg_acc(const ref_brain *r) { return r->h.bgAcc; }
int32_t ref_step(ref_brain *r) { return lif_step(&r->h); }
const int32_t *ref_last_fired(const ref_brain *r, int32_t *n) {
    *n = r->h.nFired;
    return r->h.ring + (int64_t)r->h.lastSlot * r->h.N;
}

int64_t ref_run(ref_brain *r, int32_t steps, int32_t *nfired, int32_t *raster, int64_t cap) {
    int64_t total = 0;
    for (int32_t t = 0; t < steps; t++) {
        const int32_t nf = lif_step(&r->h);
        if (nfired) nfired[t] = nf;
        if (raster) {
            const int32_t *lst = r->h.ring + (int64_t)r->h.lastSlot * r->h.N;
            for (int32_t q = 0; q < nf; q++) if (total + q < cap) raster[total + q] = lst[q];
        }
        total += nf;
    }
    return total;
}

void ref_add(ref_brain *r, int32_t which, int32_t i, double delta) {
    float *a = r->arrays[which];
    a[i] = (float)((double)a[i] + delta);
}
void ref_store(ref_brain *r, int32_t which, int32_t i, double value) { r->arrays[which][i] = (float)value; }
void ref_set_background(ref_brain *r, double rateHz, double ampMv, double dt) {
    r->h.bgEvents = (float)((double)r->h.N * rateHz * dt / 1000.0);
    r->h.bgAmp = (float)ampMv;
}
