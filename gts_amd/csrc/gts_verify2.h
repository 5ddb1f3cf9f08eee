// Query-block-stationary verification kernels (gts_verify2.hip).
#pragma once

#include "gts_index.h"

namespace gts {

struct Search;
struct VerifyArgs;

struct V2Out {
    unsigned long long *counts;
    uint64_t *cand;
    int32_t *ccount;
    int cap;
    int32_t *band_q, *band_p;
    unsigned long long *band_n;
    int64_t band_cap;
    uint64_t *res_k;
    float *res_d;
    unsigned long long *res_n;
    int64_t res_cap;
    unsigned long long *verified;
};

// leaf records sorted by (query, parent) with per-query offsets
struct V2Recs {
    int64_t n = 0;
    const int32_t *rec_q = nullptr, *rec_n = nullptr;
    const uint32_t *rec_m = nullptr;
    int64_t *roff = nullptr;
};

bool v2_supported(const gts_index *idx, int mode);
V2Recs v2_prepare(gts_index *idx, Search &S);
void v2_launch(gts_index *idx, Search &S, const V2Recs &r, int mode, const VerifyArgs &va);
// exhaustive plan: every (query, row) pair through the dense kernel
void dense_all_launch(gts_index *idx, const Search &S, const VerifyArgs &va);
// BF16 hi/lo LDS images of the batch's 64-query blocks (I->qimg); qpos null:
// only sizes the buffer
uint16_t *stage_query_image(gts_index *idx, const int32_t *qpos, int64_t nq);
// true when the exhaustive plan's launch for nq queries needs no schedule timing
bool dense_schedule_known(const gts_index *idx, int64_t nq);
// D[q][p] = d(query q, tree position piv[p]) (vector metrics)
void pivot_table(gts_index *idx, const int32_t *qpos, int64_t nq, const int32_t *piv, int64_t np, float *D);
// the same for L2 from BF16 hi x hi MFMAs (radius features), against the
// pivot rows' hi copies (ph: fragment-major, padded to 32 rows; pn2: norms)
void pivot_table_bf16(gts_index *idx, const int32_t *qpos, int64_t nq, const uint16_t *ph, const float *pn2,
                      int64_t np, float *D);
bool dense_pivot_dist(gts_index *idx, const int32_t *qpos, int64_t nq, int lvl, const int32_t *fq,
                      const int32_t *fn, int64_t nf, float *dq);

}  // namespace gts
