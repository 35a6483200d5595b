/* CPU restatement of LIFNetwork (src/lif.js:16-105).  TEST INFRASTRUCTURE ONLY.
 *
 * Values live in binary32 arrays; every expression is evaluated in binary64
 * and rounded once when stored, as JavaScript does with Float32Array.  The
 * awake list, the sleep rule and the order of the random draws follow the
 * reference line by line, so with the same mulberry32 seed this reproduces
 * its fired lists and final state exactly (pinned by tests/golden/lifnet_golden.json).
 */
#include "lifnet_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* fly-brain src/lif.js:5 */
static const double EXC_SIGN[8] = {0, 1, -1, -1, 1, 1, 1, -1};

struct lno_net {
    int32_t N;
    lno_params p;
    const uint32_t *indptr, *indices;
    uint32_t *ip, *ix;          /* own copies */
    uint8_t *nt;
    double ntSign[8];
    double *preSign;            /* NULL: ntSign[nt[j]] */
    float *w;                   /* baked weights */
    float *v, *gE, *gI, *refr, *trace, *adapt, *res, *bias, *thr, *drive;
    uint32_t *spikeCount;
    uint8_t *awake;
    int32_t *awakeList, nAwake;
    int32_t nring, head;        /* ring of nring fired lists */
    int32_t *ring, *ringCount;
    int32_t *fired, nf;         /* the last step's list */
    double t;
    uint32_t rs;                /* mulberry32 state */
    int g2set;
    double g2;
};

static inline float f32(double x) { return (float)x; }

/* Math.random as mulberry32 (scripts/lif_ensemble.mjs:9-12) */
static double rnd(lno_net *n) {
    n->rs += 0x6D2B79F5u;
    uint32_t t = (n->rs ^ (n->rs >> 15)) * (1u | n->rs);
    t = (t + (t ^ (t >> 7)) * (61u | t)) ^ t;
    return (double)(t ^ (t >> 14)) / 4294967296.0;
}

/* gauss(), src/lif.js:55-57: Marsaglia's polar method with the second value cached */
static double gauss(lno_net *n) {
    if (n->g2set) { n->g2set = 0; return n->g2; }
    double u, s, w;
    do {
        u = rnd(n) * 2 - 1;
        s = rnd(n) * 2 - 1;
        w = u * u + s * s;
    } while (w >= 1 || w == 0);
    const double m = sqrt(-2 * log(w) / w);
    n->g2 = s * m;
    n->g2set = 1;
    return u * m;
}

static double pre_sign(const lno_net *n, int32_t j) { return n->preSign ? n->preSign[j] : n->ntSign[n->nt[j]]; }

static void wake(lno_net *n, int32_t i) {   /* src/lif.js:45 */
    if (!n->awake[i]) { n->awake[i] = 1; n->awakeList[n->nAwake++] = i; }
}

/* _setupRing (src/lif.js:46): JS Math.round is floor(x + 0.5) */
static void setup_ring(lno_net *n) {
    double d = floor(n->p.delay / n->p.dt + 0.5);
    if (d < 1) d = 1;
    n->nring = (int32_t)d + 1;
    free(n->ring); free(n->ringCount);
    n->ring = (int32_t *)calloc((size_t)n->nring * n->N, sizeof(int32_t));
    n->ringCount = (int32_t *)calloc((size_t)n->nring, sizeof(int32_t));
    n->head = 0;
}

lno_net *lno_create(int32_t N, const uint32_t *indptr, const uint32_t *indices, const float *weights,
                    const uint8_t *nt, const lno_params *p, const float *inScale, const float *outScale,
                    const uint8_t *sensory, const double *ntSign, const double *preSign, uint32_t rseed) {
    lno_net *n = (lno_net *)calloc(1, sizeof *n);
    n->N = N;
    n->p = *p;
    const int64_t E = indptr[N];
    n->ip = (uint32_t *)malloc(sizeof(uint32_t) * (N + 1));
    memcpy(n->ip, indptr, sizeof(uint32_t) * (N + 1));
    n->ix = (uint32_t *)malloc(sizeof(uint32_t) * (E ? E : 1));
    memcpy(n->ix, indices, sizeof(uint32_t) * E);
    n->indptr = n->ip;
    n->indices = n->ix;
    n->nt = (uint8_t *)malloc(N ? N : 1);
    memcpy(n->nt, nt, N);
    for (int k = 0; k < 8; k++) n->ntSign[k] = ntSign ? ntSign[k] : EXC_SIGN[k];
    if (preSign) {
        n->preSign = (double *)malloc(sizeof(double) * N);
        memcpy(n->preSign, preSign, sizeof(double) * N);
    }
    /* constructor weight bake, src/lif.js:24-29 */
    n->w = (float *)malloc(sizeof(float) * (E ? E : 1));
    for (int32_t j = 0; j < N; j++) {
        const double og = outScale ? (double)outScale[j] : 1.0;
        const int keepSensory = p->preInh > 0 && pre_sign(n, j) < 0;
        for (uint32_t k = indptr[j]; k < indptr[j + 1]; k++) {
            const uint32_t q = indices[k];
            const double c = weights[k];
            const int sens = sensory ? sensory[q] : 0;
            n->w[k] = (c >= p->minSyn && (!sens || keepSensory)) ? f32(c * (inScale ? (double)inScale[q] : 1.0) * og) : 0.f;
        }
    }
    if (p->inhGain != 1) {   /* src/lif.js:30-31 */
        for (int32_t j = 0; j < N; j++)
            if (pre_sign(n, j) < 0)
                for (uint32_t k = indptr[j]; k < indptr[j + 1]; k++) n->w[k] = f32((double)n->w[k] * p->inhGain);
    }
    float **arrs[] = {&n->v, &n->gE, &n->gI, &n->refr, &n->trace, &n->adapt, &n->res, &n->bias, &n->thr, &n->drive};
    for (unsigned a = 0; a < sizeof arrs / sizeof *arrs; a++) *arrs[a] = (float *)calloc(N ? N : 1, sizeof(float));
    for (int32_t i = 0; i < N; i++) { n->v[i] = f32(p->vRest); n->res[i] = 1.f; }
    n->spikeCount = (uint32_t *)calloc(N ? N : 1, sizeof(uint32_t));
    n->awake = (uint8_t *)calloc(N ? N : 1, 1);
    n->awakeList = (int32_t *)calloc(N ? N : 1, sizeof(int32_t));
    n->fired = (int32_t *)calloc(N ? N : 1, sizeof(int32_t));
    n->rs = rseed;
    setup_ring(n);
    return n;
}

void lno_destroy(lno_net *n) {
    if (!n) return;
    free(n->ip); free(n->ix); free(n->nt); free(n->preSign); free(n->w);
    free(n->v); free(n->gE); free(n->gI); free(n->refr); free(n->trace); free(n->adapt); free(n->res);
    free(n->bias); free(n->thr); free(n->drive); free(n->spikeCount); free(n->awake); free(n->awakeList);
    free(n->ring); free(n->ringCount); free(n->fired);
    free(n);
}

/* src/lif.js:48-49 */
void lno_reset(lno_net *n) {
    for (int32_t i = 0; i < n->N; i++) {
        n->v[i] = f32(n->p.vRest);
        n->gE[i] = n->gI[i] = n->refr[i] = n->trace[i] = n->adapt[i] = 0.f;
        n->res[i] = 1.f;
        n->spikeCount[i] = 0;
        n->awake[i] = 0;
    }
    for (int32_t s = 0; s < n->nring; s++) n->ringCount[s] = 0;
    n->t = 0;
    n->nAwake = 0;
    for (int32_t i = 0; i < n->N; i++)
        if (n->drive[i] > 0 || n->bias[i] != 0) wake(n, i);
}

void lno_set_bias(lno_net *n, const int32_t *ix, int32_t k, double mv) {   /* :50 */
    for (int32_t a = 0; a < k; a++) { n->bias[ix[a]] = f32(mv); if (mv != 0) wake(n, ix[a]); }
}
void lno_set_drive(lno_net *n, const int32_t *ix, int32_t k, double rate) {   /* :51 */
    for (int32_t a = 0; a < k; a++) { n->drive[ix[a]] = f32(rate); if (rate > 0) wake(n, ix[a]); }
}
void lno_set_drive_one(lno_net *n, int32_t i, double rate) {   /* :52 */
    n->drive[i] = f32(rate);
    if (rate > 0) wake(n, i);
}
void lno_set_thr(lno_net *n, int32_t i, double mv) { n->thr[i] = f32(mv); }   /* :53 */
void lno_add_g(lno_net *n, int32_t i, double e, double ii) {   /* :54 */
    n->gE[i] = f32((double)n->gE[i] + e);
    n->gI[i] = f32((double)n->gI[i] + ii);
    wake(n, i);
}
void lno_pulse(lno_net *n, const int32_t *ix, int32_t k, double mv) {   /* :55 */
    for (int32_t a = 0; a < k; a++) { n->gE[ix[a]] = f32((double)n->gE[ix[a]] + mv); wake(n, ix[a]); }
}

/* step(), src/lif.js:59-105 */
int32_t lno_step(lno_net *n) {
    const lno_params *P = &n->p;
    const int32_t N = n->N;
    const double dt = P->dt, vRest = P->vRest, vThresh = P->vThresh, vReset = P->vReset;
    const double dA = exp(-dt / P->adaptTau), kRec = dt / P->depTau;
    const int coba = P->coba != 0;
    const double eExc = P->eExc, eInh = P->eInh, cE = 1 / (eExc - vRest), cI = 1 / (vRest - eInh);
    const double preInh = P->preInh;
    const double dE = exp(-dt / P->tauSyn), dTr = exp(-dt / P->traceTau), dtS = dt / 1000,
                 nA = P->noise * sqrt(dt), kM = dt / P->tauM;
    const int32_t *arriving = n->ring + (int64_t)n->head * N;
    const int32_t na = n->ringCount[n->head];
    int32_t nA0 = n->nAwake;
    for (int32_t k = 0; k < na; k++) {
        const int32_t pre = arriving[k];
        const double pre_g = preInh > 0 && n->drive[pre] > 0
                                 ? 1 / (1 + preInh * fmax(0, -(double)n->gI[pre])) : 1;
        const double sign = pre_sign(n, pre) * P->wSyn * (double)n->res[pre] * pre_g;
        if (sign == 0) continue;
        n->res[pre] = f32((double)n->res[pre] - P->depU * (double)n->res[pre]);
        float *g = sign > 0 ? n->gE : n->gI;
        for (uint32_t j = n->indptr[pre]; j < n->indptr[pre + 1]; j++) {
            const uint32_t q = n->indices[j];
            const double w = n->w[j];
            if (w == 0) continue;
            g[q] = f32((double)g[q] + w * sign);
            if (!n->awake[q]) { n->awake[q] = 1; n->awakeList[nA0++] = (int32_t)q; }
        }
    }
    n->ringCount[n->head] = 0;
    int32_t nf = 0;
    const int full = nA > 0;
    const int32_t count = full ? N : nA0;
    int32_t keep = 0;
    for (int32_t c = 0; c < count; c++) {
        const int32_t i = full ? c : n->awakeList[c];
        double vi = n->v[i];
        if (n->refr[i] > 0) {
            n->refr[i] = f32((double)n->refr[i] - dt);
            vi = vReset;
        } else if (n->drive[i] > 0 && rnd(n) < (double)n->drive[i] * dtS) {
            vi = vReset;
            n->refr[i] = f32(P->tRef);
            n->fired[nf++] = i;
            n->spikeCount[i]++;
            n->trace[i] = 1.f;
            n->adapt[i] = f32((double)n->adapt[i] + P->adaptInc);
        } else {
            const double gEi = n->gE[i], gIi = n->gI[i], bi = n->bias[i];
            if (coba) vi += (vRest - vi + gEi * (eExc - vi) * cE + gIi * (vi - eInh) * cI + bi) * kM;
            else vi += (vRest - vi + gEi + gIi + bi) * kM;
            if (nA > 0) vi += nA * gauss(n);
            if (vi >= vThresh + (double)n->adapt[i] + (double)n->thr[i]) {
                vi = vReset;
                n->refr[i] = f32(P->tRef);
                n->fired[nf++] = i;
                n->spikeCount[i]++;
                n->trace[i] = 1.f;
                n->adapt[i] = f32((double)n->adapt[i] + P->adaptInc);
            }
        }
        const double g1 = (double)n->gE[i] * dE, g2 = (double)n->gI[i] * dE, tr = (double)n->trace[i] * dTr,
                     ad = (double)n->adapt[i] * dA, rs = (double)n->res[i] + (1 - (double)n->res[i]) * kRec;
        if (!full && n->refr[i] <= 0 && n->drive[i] == 0 && n->bias[i] == 0 && vi - vRest < 1e-3 &&
            vRest - vi < 1e-3 && g1 < 1e-4 && -g2 < 1e-4 && tr < 2e-3 && ad < 1e-3 && rs > 0.999) {
            n->v[i] = f32(vRest);   /* back to rest: sleep */
            n->gE[i] = n->gI[i] = n->trace[i] = n->adapt[i] = 0.f;
            n->res[i] = 1.f;
            n->awake[i] = 0;
        } else {
            n->v[i] = f32(vi);
            n->gE[i] = f32(g1);
            n->gI[i] = f32(g2);
            n->trace[i] = f32(tr);
            n->adapt[i] = f32(ad);
            n->res[i] = f32(rs);
            if (!full) n->awakeList[keep++] = i;
        }
    }
    if (!full) n->nAwake = keep;
    else {
        n->nAwake = 0;
        for (int32_t i = 0; i < N; i++)
            if (n->awake[i]) n->awakeList[n->nAwake++] = i;
    }
    const int32_t slot = (n->head + n->nring - 1) % n->nring;
    memcpy(n->ring + (int64_t)slot * N, n->fired, sizeof(int32_t) * nf);
    n->ringCount[slot] = nf;
    n->nf = nf;
    n->head = (n->head + 1) % n->nring;
    n->t += dt;
    return nf;
}

const int32_t *lno_fired(const lno_net *n) { return n->fired; }

float *lno_array(lno_net *n, int32_t which) {
    switch (which) {
    case LNO_V: return n->v;
    case LNO_GE: return n->gE;
    case LNO_GI: return n->gI;
    case LNO_REFR: return n->refr;
    case LNO_TRACE: return n->trace;
    case LNO_ADAPT: return n->adapt;
    case LNO_RES: return n->res;
    case LNO_BIAS: return n->bias;
    case LNO_THR: return n->thr;
    case LNO_DRIVE: return n->drive;
    case LNO_WEIGHTS: return n->w;
    default: return 0;
    }
}

const uint32_t *lno_spike_count(const lno_net *n) { return n->spikeCount; }
int32_t lno_n_awake(const lno_net *n) { return n->nAwake; }
double lno_time(const lno_net *n) { return n->t; }
int32_t lno_nslots(const lno_net *n) { return n->nring; }

int64_t lno_run(lno_net *n, int32_t steps, int32_t *nf) {
    int64_t tot = 0;
    for (int32_t s = 0; s < steps; s++) {
        const int32_t k = lno_step(n);
        if (nf) nf[s] = k;
        tot += k;
    }
    return tot;
}
