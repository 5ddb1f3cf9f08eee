// Cell scan of the learned kNN form (gts_cells.hip): a Voronoi partition of
// the rows over sampled pivot rows, rows copied in cell order, and per batch
// the (query block, 32-row chunk) pairs that the triangle inequality cannot
// exclude, verified by the BF16 MFMA kernel.
#pragma once

#include "gts_index.h"

namespace gts {

struct Search;
struct VerifyArgs;

// L2 rows in the BF16 form and no deletions pending: the cell scan applies
bool cells_supported(const gts_index *idx);
// builds the partition and the cell-ordered copies once per tree (no-op when built)
void cells_build(gts_index *idx);
void cells_release(gts_index *idx);
// BF16 split copies (fragment-major, rows padded with zeros to npad) and norms
// of the rows at tree positions pos[0..n): hi/lo hold npad * dk, n2 holds n
void gather_split_rows(gts_index *idx, const int32_t *pos, int64_t n, int64_t npad, uint16_t *hi, uint16_t *lo,
                       float *n2);
// orders a batch by the cell-order row of each query (qpos, kq, bnd, orig permuted in place)
void cells_sort_queries(gts_index *idx, int32_t *qpos, int32_t *kq, float *bnd, int32_t *orig, int64_t nq);
// what is already in place for cells_launch
enum {
    CELLS_STAGED = 1,    // the query image of S.qpos (stage_query_image)
    CELLS_CLEARED = 2,   // the first segment's live-chunk bitmap (cells_prepare_batch)
};
// every pair of the batch that the cell bounds keep, through the MFMA kernel;
// ev0/ev1 (optional) bracket the verification kernels, as external records
// when `external` (graph capture); `ready`: CELLS_* flags; `rerun`: a rerun
// batch (its own kernel instance)
void cells_launch(gts_index *idx, const Search &S, const VerifyArgs &va, hipEvent_t ev0, hipEvent_t ev1,
                  bool external, int ready, bool rerun = false);
// a batch of query IDs (k from ks, or kfix) prepared in the scan order:
// qpos, kq, orig (batch index), bnd (+inf) ordered as cells_sort_queries
// orders them; ccount[0..nq), the pass scalars scal[8..14] (scal8: 14 words) and the first segment's
// live-chunk bitmap cleared (two kernels).  IDs outside [0, n) and k outside
// [1, kcap] set *argerr (published with the pass) and are clamped.  Needs the
// rank array all zero (cells_ranks_ready, outside graph capture) and leaves
// it so.
void cells_prepare_batch(gts_index *idx, const int32_t *qids, const int32_t *ks, int kfix, int kcap,
                         int32_t *argerr, int64_t nq, int32_t *qpos, int32_t *kq, int32_t *orig, float *bnd,
                         int32_t *ccount, int32_t *scal8);
// sizes the rank array for nq queries and clears it unless it is known clear
// (a stream fill, not for graph capture)
void cells_ranks_ready(gts_index *idx, int64_t nq);
// the live-chunk bitmap of a batch of nq queries' first segment (*nwords
// words), for callers that clear it themselves (CELLS_CLEARED)
uint32_t *cells_bitmap(gts_index *idx, int64_t nq, int64_t *nwords);
// pre-sizes every scratch buffer of cells_launch for nq queries and mixes their
// addresses into `mix` (graph capture allocates nothing)
void cells_reserve(gts_index *idx, int64_t nq, void (*mix)(void *, uint64_t), void *ctx);

}  // namespace gts
