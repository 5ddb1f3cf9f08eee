// Query drivers behind the C ABI (gts_query.hip, gts_update.hip, gts_learned.hip).
#pragma once

#include "gts_index.h"

namespace gts {

struct RangeOut {
    std::vector<int64_t> counts;
    std::vector<int32_t> ids;    // CSR by query, ascending ID
    std::vector<float> dists;
};

void range_run(gts_index *idx, const int32_t *qids_host, int64_t nq, float r, bool want_ids,
               RangeOut *out);
// range counts of device query positions under snapshot visibility (update
// path): position p is hidden from query q when d_death[p] < d_qop[q]
void range_count_dev(gts_index *idx, const int32_t *d_qpos, const int32_t *d_qop, int64_t nq, float r,
                     const int32_t *d_death, int64_t *h_counts);
void knn_run(gts_index *idx, const int32_t *d_qids, int64_t nq, const int32_t *d_ks, int kfix,
             int kstride, int32_t *d_ids, float *d_dists, float *d_kth, int learned, int verify);
void update_run(gts_index *idx, const gts_update_op *ops, int64_t nops, float r, int cache_cap,
                const float *fresh, int64_t nfresh, int64_t *counts, int64_t cap,
                int64_t *nqueries);
float learned_train(gts_index *idx, const int32_t *train_qids, int64_t ntrain, int kmax,
                    int epochs, float lr, uint64_t seed);
// radius of each query into bnd; true when the BF16 query image of
// qpos[0..nq) (stage_query_image) is left staged
bool learned_predict(gts_index *idx, const int32_t *qpos, const int32_t *kq, int64_t nq,
                     float *bnd);
void learned_apply_quantile(LearnedModel &m);
void release_scratch(gts_index *idx);
// adds the last kNN call's event timings (pend_timing bits) to its stats
void finalize_timing(gts_index *idx, int bits = 7);
int mlp_num_params();
float mlp_loss_grad(gts_index *idx, const float *X, const float *t, int64_t nb, const float *w,
                    float *grad);
// the same on device buffers, allocating nothing (a fixed per-GPU array)
void mlp_loss_grad_noalloc(gts_index *idx, const float *X, const float *t, int64_t nb, const float *w,
                           float *grad, float *loss);

}  // namespace gts
