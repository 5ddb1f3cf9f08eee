// Host-side driver of the two-stage search kernels (gts_search.hip).
#pragma once

#include "gts_index.h"

namespace gts {

struct VerifyArgs {
    int64_t *counts = nullptr;    // range: per-query counts
    uint64_t *cand = nullptr;     // kNN: candidates [nq * cap]
    int32_t *ccount = nullptr;    // kNN: candidates per query
    int cap = 0;
    int32_t *band_q = nullptr, *band_p = nullptr;
    int64_t *band_n = nullptr;
    int64_t band_cap = 0;
    uint64_t *res_k = nullptr;    // range IDs: (query << 32 | ID)
    float *res_d = nullptr;
    int64_t *res_n = nullptr;
    int64_t res_cap = 0;
    int64_t *verified = nullptr;
    int64_t *live = nullptr;      // cell scan: (query block, chunk) pairs verified
    int64_t *ncand = nullptr;     // cell scan: candidates written
};

// One batch of queries through the pipeline.  Per-query device arrays are
// indexed by the batch-local query index.
struct Search {
    gts_index *idx = nullptr;
    int64_t nq = 0;
    const int32_t *qpos = nullptr;  // query object positions in tree order
    float *bnd = nullptr;           // per-query bound: radius, or kNN bound
    const int32_t *kq = nullptr;    // per-query k (kNN)
    int64_t nrec = 0, nslots = 0, ntiles = 0;
    int rec_levels = 0;                    // tree levels whose parents produced records
    int32_t rec_pmin = 0, rec_pmax = -1;   // node-ID span of those parents
    bool grouped = false, v2_ready = false;
    bool mfma_keys = false;                // kNN keys came from the FP32 MFMA tile kernel
    gts_stats stats{};

    void expand(bool knn);
    void group();
    void verify(int mode, VerifyArgs &va);
    void verify_all(VerifyArgs &va);   // kNN, exhaustive plan: no expand() before it
    void band(int mode, VerifyArgs &va, int64_t nb);
    // fail/ctr (optional): the queries left with a nonzero status appended to
    // fail[] (count in ctr[0]); with count_candidates, the candidate total
    // added to ctr[1]
    void topk(VerifyArgs &va, int check_radius, const int32_t *orig, int kstride, int32_t *ids,
              float *dists, float *kth, int32_t *status, int32_t *fail = nullptr,
              unsigned long long *ctr = nullptr, bool count_candidates = true);

  private:
    void ensure_records(int64_t need);
};

int topk_cap_limit();

// Pivot positions of the existing split nodes at the deepest level above the
// leaves with at most maxp of them, or at the shallowest one with at least
// minp of them when that comes first (the nodes of one level hold disjoint
// ranges, so the pivots are distinct objects); *level receives the level
// (-1: none).
std::vector<int32_t> level_pivots(gts_index *idx, int maxp, int *level, int minp = INT32_MAX);

// Radius source GTS_RADIUS_PIVOTS: bnd[q] = the kq[q]-th smallest FP32
// distance from query q to the level pivots (at most RADIUS_MAXP), rounded up
// by the FP32 error bound -- updateDisK's bound (search_v2.cuh:755-777) over
// one flat pivot level, with no network; +inf where kq[q] exceeds the pivot
// count.  Stream operations only once pivot_radius_buffers(idx, nq) has run.
constexpr int RADIUS_MAXP = 4096;
void pivot_radius_prepare(gts_index *idx);
void pivot_radius_buffers(gts_index *idx, int64_t nq);
void pivot_radius(gts_index *idx, const int32_t *qpos, const int32_t *kq, int64_t nq, float *bnd);

}  // namespace gts
