// fb_batch.h -- device-side description of an ensemble (shared by the
// per-step ensemble kernels and the ensemble lookahead kernel).
//
// Layout in HBM: per-member state [N][B] (member b of neuron i at i*B + b, so
// a wave with lane = member reads one neuron of every member in one
// transaction); the delay line as per-slot spike and forced masks (bit b =
// member b) with a list of neurons whose mask is nonzero; accumulators
// [2N][B] with the ordered-sum chain pool of the forward engine.
#pragma once
#include "fb_internal.h"

namespace fb {

// One arrival of the chain pool: the later same-step arrivals of a (sign
// class, target, member) hang off chain[] in arbitrary order.
struct CollEntry {
    uint64_t key;
    float addend;
    int32_t next;
};
enum { FB_DEVERR_COLL_OVERFLOW = FB_DEVERR_RUN_OVERFLOW };

// Sum of the same-step arrivals on one (sign class, target, member) in
// arrival order: the first sits in (key, add), the others on the chain.
// Insertion-sorts up to 32 of them; beyond that, repeated minimum search.
__device__ inline float ordered_sum_chain(float g, int32_t count, uint64_t key0, float add0, int32_t head,
                                          const CollEntry *coll) {
    if (count <= 32) {
        uint64_t ks[32];
        float as[32];
        int n = 0;
        ks[n] = key0; as[n] = add0; n++;
        for (int32_t p = head; p >= 0 && n < 32; p = coll[p].next) {
            uint64_t kk = coll[p].key;
            float aa = coll[p].addend;
            int m = n++;
            while (m > 0 && ks[m - 1] > kk) { ks[m] = ks[m - 1]; as[m] = as[m - 1]; m--; }
            ks[m] = kk; as[m] = aa;
        }
        for (int m = 0; m < n; m++) g = g + as[m];
        return g;
    }
    uint64_t last = 0;
    bool first = true;
    for (int32_t c = 0; c < count; c++) {
        uint64_t best = ~0ull;
        float ba = 0.f;
        if ((first || key0 > last) && key0 < best) { best = key0; ba = add0; }
        for (int32_t p = head; p >= 0; p = coll[p].next) {
            uint64_t kk = coll[p].key;
            if ((first || kk > last) && kk < best) { best = kk; ba = coll[p].addend; }
        }
        g = g + ba;
        last = best;
        first = false;
    }
    return g;
}

constexpr int BMAX = 64;
constexpr int DRV_WG = 4;          // workgroups of the multi-workgroup draws
constexpr int LA_NSHAPES = 4;      // regular-block shapes of the ensemble lookahead engine

struct BatchDev {
    int32_t N, B, nslots, nDriven, nTypes;
    // adapt stays exactly 0 without adaptInc, res exactly 1 without depU, and
    // there are no background hits without bgRate: those arrays are then
    // neither read nor written
    int32_t liveAdapt, liveRes, liveBg;
    Consts k;
    DevGraph g;
    // [N][B]
    float *v, *gE, *gI, *refr, *trace, *adapt, *res, *bias, *thr;
    uint32_t *spikeCount;
    float *drive;                  // [N], shared by every member
    // lane mode (fb_batch_set_lanes): signGain delivers W * f32(sign * gain)
    // (the per-lane type gain scales the presynaptic sign, kernel.bend's
    // lanes) instead of f32(count * gain) * sign; laneDrive [N][B] gives each
    // member its own rates (null: drive[N] for all)
    int32_t signGain;
    float *laneDrive;
    // the per-lane pool loop (fb_batch_set_pool_feedback); poolOf null when off
    const uint8_t *poolOf;
    const int32_t *loopTargets, *loopDelay;   // [nTargets], [B]
    const int32_t *loopCells;      // [loopNCells] the neurons that belong to a pool
    uint32_t *loopHist;            // [B][loopRing]
    uint16_t *loopRec;             // [B][loopRecCap][loopPools]
    int32_t *loopState;            // [0] steps since set, [1] records written
    int32_t loopPools, loopNTargets, loopRing, loopRecCap, loopNCells;
    float loopBase, loopGain;
    const int32_t *driven;         // [nDriven]
    const int32_t *drvPos;         // [N]
    const int32_t *typeOf;         // [N]
    const float *gain;             // [B][nTypes]
    const float *gainT;            // [nTypes][B], the same values by type
    uint32_t *rng;                 // [B]
    // the draws over several workgroups (no background events): each wave's
    // non-refractory count per lane, and the stream state after the step,
    // handed to kb_membrane
    int32_t *drvPart;              // [DRV_WG * 16][64]
    uint32_t *rngNext;             // [B]
    int32_t multiDrv;              // this step's draws took the multi-workgroup path
    float *bgAcc;                  // [1] (the same for every member)
    int32_t *steps;                // [B] run length of each member
    // delay line: per slot, spike and forced masks per neuron and a list of
    // neurons with a nonzero mask
    uint64_t *mask, *fmask;        // [nslots][N]
    int32_t *frank;                // [nslots][N] driven-list position of a forced spike when it was emitted
                                   // (its arrival rank; the list may change while it is in flight)
    int32_t *active, *nActive;     // [nslots][N], [nslots]
    // accumulators [2N][B] and the chain pool
    int32_t *cnt, *chain;
    float *add;
    uint64_t *key;
    CollEntry *coll;
    int64_t collCap;
    int32_t *ncoll;
    int32_t *bgCnt;                // [N][B]
    const uint32_t *leap;
    int32_t *err;
    unsigned long long *counters;  // [0] edge events, [1] synapse events, [2] spikes
};

}  // namespace fb
