/* CPU restatement of the reference's JS twin LIFNetwork (src/lif.js:16-105).
 * TEST INFRASTRUCTURE ONLY: the checker for the LIFNetwork construction path
 * and the statistical comparison with the HIP engine.
 *
 * JavaScript semantics: arithmetic in binary64, every store into a
 * Float32Array rounded to binary32, Math.random replaced by mulberry32 (the
 * stream the reference's ensemble scripts install, scripts/lif_ensemble.mjs:8-12).
 */
#ifndef LIFNET_ORACLE_H
#define LIFNET_ORACLE_H
#include <stdint.h>

typedef struct lno_params {
    double dt, vRest, vThresh, vReset, tauM, tauSyn, tRef, delay, wSyn, noise, traceTau;
    double adaptInc, adaptTau, depU, depTau, minSyn, preInh, coba, eExc, eInh, inhGain;
} lno_params;

typedef struct lno_net lno_net;

enum { LNO_V, LNO_GE, LNO_GI, LNO_REFR, LNO_TRACE, LNO_ADAPT, LNO_RES, LNO_BIAS, LNO_THR, LNO_DRIVE, LNO_WEIGHTS };

/* inScale, outScale, sensory, ntSign (8 entries), preSign may be NULL */
lno_net *lno_create(int32_t N, const uint32_t *indptr, const uint32_t *indices, const float *weights,
                    const uint8_t *nt, const lno_params *p, const float *inScale, const float *outScale,
                    const uint8_t *sensory, const double *ntSign, const double *preSign, uint32_t rseed);
void lno_destroy(lno_net *n);
void lno_reset(lno_net *n);
void lno_set_bias(lno_net *n, const int32_t *ix, int32_t k, double mv);
void lno_set_drive(lno_net *n, const int32_t *ix, int32_t k, double rate);
void lno_set_drive_one(lno_net *n, int32_t i, double rate);
void lno_set_thr(lno_net *n, int32_t i, double mv);
void lno_add_g(lno_net *n, int32_t i, double e, double ii);
void lno_pulse(lno_net *n, const int32_t *ix, int32_t k, double mv);
int32_t lno_step(lno_net *n);
const int32_t *lno_fired(const lno_net *n);
float *lno_array(lno_net *n, int32_t which);
const uint32_t *lno_spike_count(const lno_net *n);
int32_t lno_n_awake(const lno_net *n);
double lno_time(const lno_net *n);
int32_t lno_nslots(const lno_net *n);
/* steps without keeping the lists; per-step spike counts into nf (may be NULL) */
int64_t lno_run(lno_net *n, int32_t steps, int32_t *nf);

#endif
