/*
 * lif_oracle.h -- declarations for the CPU restatement of the reference LIF
 * step (lif_oracle.c).  TEST INFRASTRUCTURE ONLY: tests/, smoke() and the
 * bench's CPU baseline use it as the checker; the product never links it.
 */
#ifndef FLYBRAIN_LIF_ORACLE_H
#define FLYBRAIN_LIF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Model parameters as the JS wrapper holds them (f64), in the order of the
 * product's fb_params struct (include/flybrain.h). */
typedef struct {
    double dt, vRest, vThresh, vReset, tauM, tauSyn, tRef, delay;
    double traceTau, adaptInc, adaptTau, depU, depTau;
    double coba, eExc, eInh;
    double bgRate, bgAmp, preInh;
} fbo_params;

/* f32 header constants (lifwasm.js:58-67) */
typedef struct {
    float dt, vRest, vThresh, vReset, tRef, adaptInc, depU;
    float dE, dTr, dA, kRec, kM, dtS, eExc, eInh, cE, cI;
    float bgEvents, bgAmp, preInh;
    int32_t coba;
} fbo_consts;

enum {
    FBO_V = 0, FBO_GE, FBO_GI, FBO_REFR, FBO_TRACE, FBO_ADAPT, FBO_RES,
    FBO_BIAS, FBO_THR, FBO_DRIVE, FBO_NARRAYS
};

typedef struct fbo_brain fbo_brain;

uint32_t fbo_xorshift32(uint32_t *state);
uint32_t fbo_seed_hash(double seed);
int32_t fbo_nslots(double delay, double dt);
void fbo_constants(const fbo_params *p, int32_t N, fbo_consts *c);

fbo_brain *fbo_create(int32_t N, int64_t E, const uint32_t *indptr,
                      const uint32_t *indices, const float *W,
                      const float *sign, const fbo_params *p, double seed);
void fbo_destroy(fbo_brain *b);
void fbo_reset(fbo_brain *b);
void fbo_set_driven(fbo_brain *b, const int32_t *list, int32_t n);
float *fbo_array(fbo_brain *b, int32_t which);
uint32_t *fbo_spike_count(fbo_brain *b);
uint32_t fbo_rng_state(const fbo_brain *b);
/* the xorshift32 state itself (a cord-ensemble lane starts from a raw state) */
void fbo_set_rng(fbo_brain *b, uint32_t rng);
int32_t fbo_head(const fbo_brain *b);
int32_t fbo_nslots_of(const fbo_brain *b);
float fbo_bg_acc(const fbo_brain *b);
const int32_t *fbo_last_fired(const fbo_brain *b, int32_t *n);
int32_t fbo_step(fbo_brain *b);
int64_t fbo_run(fbo_brain *b, int32_t steps, int32_t *nfired, int32_t *raster, int64_t cap);
void fbo_add(fbo_brain *b, int32_t which, int32_t i, double delta);
void fbo_store(fbo_brain *b, int32_t which, int32_t i, double value);
void fbo_set_background(fbo_brain *b, double rateHz, double ampMv, double dt);

/* The flyvis optic-lobe rate network's step (src/wasm/lif.c:82-93):
 * acc = sum over sources j with v_j > 0 (ascending j, each row in order) of
 * weight * v_j into acc[target]; then v += kdt * (-v + bias + acc + x). */
void fbo_fv_step(int32_t N, const float *bias, const float *kdt, const int32_t *indptr, const int32_t *target,
                 const float *weight, float *v, float *acc, const float *x);
/* One reverse step of FlyVisDiff.backward (src/flyvisdiff.js:154-182) at
 * step t: vPre / uAt are the step's pre-update activity and integrator
 * argument, dLdV its loss term; lv / lacc are the carried adjoints; gw / gb /
 * gk accumulate; dLdLum (may be NULL) gets lacc of every input node
 * inNode[q] added to its column inCol[q], in q order.  JavaScript semantics:
 * binary64 arithmetic, binary32 stores. */
void fbo_fv_back_step(int32_t N, const int32_t *indptr, const int32_t *target, const float *weight, const float *kdt,
                      const float *vPre, const float *uAt, const float *dLdV, float *lv, float *lacc, float *gw,
                      float *gb, float *gk, float *dLdLum, const int32_t *inNode, const int32_t *inCol, int32_t nIn);

#ifdef __cplusplus
}
#endif
#endif
