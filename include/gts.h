/*
 * gts.h -- C ABI of the MI355X (gfx950) GTS pivot-tree index.
 *
 * Plain C types only: pointers, sizes, status codes.  Every entry point returns
 * GTS_OK or an error code; gts_last_error() gives the message of the last
 * failure on the calling thread.  Handles are independent: calls on one handle
 * are serialised by the caller, calls on different handles may run
 * concurrently.  Each index owns one HIP stream on the device it was built on.
 *
 * Each declaration names the reference routine it replaces (paths under
 * "Source Code/GTS/" of ZJU-DAILY/GTS).  The reference has no library surface:
 * its entry points are header functions called from main.cu with global
 * outputs, which these calls turn into caller-owned buffers.
 */
#ifndef GTS_H
#define GTS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (the reference calls exit() instead: tree.cuh:19-29, file.cuh:31-35) */
#define GTS_OK 0
#define GTS_E_IO 1        /* file cannot be opened or read                  */
#define GTS_E_FORMAT 2    /* malformed header or row                        */
#define GTS_E_ARG 3       /* bad argument (IDs out of range, k < 1, ...)    */
#define GTS_E_OOM 4       /* device or host allocation failed               */
#define GTS_E_DEVICE 5    /* HIP runtime error                              */
#define GTS_E_CAPACITY 6  /* a caller buffer is too small (size is reported) */

/* largest k of a kNN call: the top-k's per-query candidate capacity (the
 * reference has no limit; it sorts every candidate, search_v2.cuh:1160) */
#define GTS_MAX_K 4096

/* metric codes, identical to the reference's data_info[2] (tree.cuh:81-154) */
#define GTS_METRIC_LINF 0
#define GTS_METRIC_L1 1
#define GTS_METRIC_L2 2
#define GTS_METRIC_ANGULAR 5
#define GTS_METRIC_EDIT 6

/* verification forms for L2 kNN queries; both return the exact result sets.
 * FP32: FP32 distances.  BF16X3: products hi.hi + hi.lo + lo.hi of BF16 hi/lo
 * splits on the matrix cores (the dense kernel gates tiles with hi.hi alone
 * first), FP64 re-ranking of every candidate whose error band reaches the
 * k-th distance. */
#define GTS_VERIFY_FP32 0
#define GTS_VERIFY_BF16X3 1

/* plans of a learned-radius kNN batch in the BF16X3 form (gts_set_knn_plan):
 * AUTO expands a 256-query sample through the tree and skips the pruning when
 * it would keep the dense regime's (query, row) pairs anyway, verifying
 * through the cell scan; PRUNED always prunes level by level; EXHAUSTIVE
 * verifies every (query, object) pair; CELLS verifies every pair that the
 * cell bounds keep (a Voronoi partition over pivot rows, rows copied in cell
 * order; see gts_set_cell_count).  For learned calls over L2, L1 or Linf
 * rows of at most 8 dimensions with kstride <= 64, AUTO and PERSISTENT run
 * the batch as one persistent launch (each wave walks one query's tree
 * depth-first; schedule autotuned per batch shape; repeated batches replay
 * a captured graph); PERSISTENT fails with GTS_E_ARG for other calls. */
#define GTS_PLAN_AUTO 0
#define GTS_PLAN_PRUNED 1
#define GTS_PLAN_EXHAUSTIVE 2
#define GTS_PLAN_CELLS 3
#define GTS_PLAN_PERSISTENT 4

/* radius sources of learned kNN calls (gts_set_radius_source) */
#define GTS_RADIUS_NETWORK 0
#define GTS_RADIUS_PIVOTS 1

typedef struct gts_dataset gts_dataset;
typedef struct gts_index gts_index;

typedef struct {
    int order;     /* fan-out T, 2..32           (TREE_ORDER, tree.cuh:31)       */
    int max_size;  /* leaf capacity              (MAX_SIZE,  tree.cuh:32)       */
    int max_h;     /* levels; 0 = as many as needed (MAX_H, tree.cuh:33)        */
    int device;    /* HIP device ordinal                                        */
} gts_build_params;

/* update op (UO, update.cuh:18-22): flag 0 insert (id = physical row of the
 * current data), 1 delete (id = logical ID), 2 range query (id = physical row),
 * 3 insert of a fresh row (id = index into the rows given to gts_update). */
typedef struct {
    int32_t flag;
    int32_t id;
} gts_update_op;

const char *gts_last_error(void);
int gts_version(void);

/* ---- datasets (host memory) ---------------------------------------------- */

/* load (file.cuh:28-105): header "dim N metric", then N rows of floats, or N
 * raw strings for metric 6. */
int gts_dataset_load(const char *path, gts_dataset **out);
/* the same from memory: rows [n, dim] FP32 row-major */
int gts_dataset_from_vectors(const float *rows, int64_t n, int dim, int metric,
                             gts_dataset **out);
/* strings packed at `stride` bytes with lengths `lens` (metric 6) */
int gts_dataset_from_strings(const uint8_t *chars, const int32_t *lens, int64_t n,
                             int stride, gts_dataset **out);
int gts_dataset_info(const gts_dataset *ds, int64_t *n, int *dim, int *metric);
/* copy of the FP32 rows (n * dim floats) */
int gts_dataset_rows(const gts_dataset *ds, float *out, int64_t cap_floats);
void gts_dataset_free(gts_dataset *ds);

/* loadQuery (file.cuh:112-151): first line the count, then one ID per line.
 * Fills up to `cap` IDs into `ids` and sets *nq (the count in the file). */
int gts_load_queries(const char *path, int32_t *ids, int64_t cap, int64_t *nq);
/* loadUpdate (update.cuh:43-95): count, then "flag id" per line. */
int gts_load_updates(const char *path, gts_update_op *ops, int64_t cap, int64_t *nops);

/* ---- index ------------------------------------------------------------- */

/* indexConstru (tree.cuh:332-393): pivot distances, per-node sort, T-way
 * equal-count split, level by level, on the GPU. */
int gts_build(const gts_dataset *ds, const gts_build_params *params, gts_index **out);
void gts_index_free(gts_index *idx);
/* tree height (tree_h), nodes in BFS numbering, objects */
int gts_index_info(const gts_index *idx, int *height, int64_t *num_nodes, int64_t *n);
/* tree export for checks: perm[n] (position -> object ID); per node: lid,
 * size, pivot object (-1 unless split), min_dis, flags (1 exists, 2 leaf). */
int gts_index_export(const gts_index *idx, int32_t *perm, int32_t *lid, int32_t *size,
                     int32_t *pivot, float *min_dis, int32_t *flags);

/* ---- queries: IDs and outputs in host memory ---------------------------- */

/* searchIndexRnnV2 (search_v2.cuh:780-972): counts[nq] of objects within r.
 * If `ids` is non-NULL, offsets[nq + 1] (CSR) and up to `cap` (ID, distance)
 * pairs ascending by ID per query are written; GTS_E_CAPACITY reports the
 * needed size in offsets[nq]. */
int gts_range_query(gts_index *idx, const int32_t *qids, int64_t nq, float r, int64_t *counts,
                    int64_t *offsets, int32_t *ids, float *dists, int64_t cap);

/* searchIndexKnnV2 (search_v2.cuh:975-1228): the k nearest objects of each
 * query, ascending by (distance, ID): ids/dists [nq * k]; kth[nq] is the k-th
 * distance (the reference's res_dis output, main.cu:101-104).  Any output
 * pointer may be NULL.  1 <= k <= GTS_MAX_K; any number of objects tied at
 * the k-th distance is resolved (by ID). */
int gts_knn_query(gts_index *idx, const int32_t *qids, int64_t nq, int k, int32_t *ids,
                  float *dists, float *kth);

/* The same with one k per query (uneven k); rows of ids/dists have stride
 * `kstride` >= max k, unused slots get -1 / +inf. */
int gts_knn_query_var(gts_index *idx, const int32_t *qids, int64_t nq, const int32_t *ks,
                      int kstride, int32_t *ids, float *dists, float *kth);

/* saveK (file.cuh:154-176): per query one line of its k IDs ("%d ") to
 * `ids_path` and of its k distances ("%f ") to `dists_path` (either may be
 * NULL), rows of ids/dists at stride k; the files are truncated. */
int gts_save_knn(const char *ids_path, const char *dists_path, const int32_t *ids, const float *dists,
                 int k, int64_t nq);
/* saveR (file.cuh:179-199): the same for range results in CSR form (query
 * i's pairs at offsets[i] .. offsets[i] + counts[i]); the files are appended
 * to, as the reference does. */
int gts_save_range(const char *ids_path, const char *dists_path, const int32_t *ids, const float *dists,
                   const int64_t *counts, const int64_t *offsets, int64_t nq);

/* updateIndexRnn (update.cuh:449-638): applies the ops in order, with a cache
 * list of `cache_cap` inserts (MAX_IN_SIZE, update.cuh:16) and a rebuild when
 * it fills; each query op (flag 2) writes its count to counts[]; `fresh`
 * holds nfresh rows for flag-3 inserts (may be NULL).  *nqueries receives the
 * number of query ops. */
int gts_update(gts_index *idx, const gts_update_op *ops, int64_t nops, float r, int cache_cap,
               const float *fresh, int64_t nfresh, int64_t *counts, int64_t cap,
               int64_t *nqueries);

/* ---- queries: IDs and outputs in device memory (no host copies) --------- */

/* kNN with query IDs and outputs already on the index's device; runs on the
 * index stream and returns after the results are complete.  `learned` = 1
 * uses the trained radius network (gts_learned_train) for one tight range
 * pass with exact radius doubling; `verify` selects GTS_VERIFY_FP32 or
 * GTS_VERIFY_BF16X3 (L2 only).  d_qids, d_ids, d_dists and d_kth are
 * required.  The device inputs are checked on the device: a query ID outside
 * [0, n) or a k (d_ks[i], or kfix) outside [1, min(kstride, GTS_MAX_K)] fails
 * the call with GTS_E_ARG (the outputs are then undefined, and nothing is
 * read or written out of bounds). */
int gts_knn_query_dev(gts_index *idx, const int32_t *d_qids, int64_t nq, const int32_t *d_ks,
                      int kfix, int kstride, int32_t *d_ids, float *d_dists, float *d_kth,
                      int learned, int verify);

/* Plan of later learned BF16X3 kNN calls on this index (GTS_PLAN_*; no
 * counterpart in the reference).  EXHAUSTIVE applies to learned L2 calls in the
 * BF16X3 form at padded dimension <= 512; other calls fail with GTS_E_ARG
 * while it is set.  Results are identical under every plan. */
int gts_set_knn_plan(gts_index *idx, int plan);

/* Matrix cores for the FP32 form (no counterpart in the reference): on (1),
 * L2 range counts and kNN candidates of GTS_VERIFY_FP32 calls come from FP32
 * MFMA tiles (d^2 = |q|^2 + |x|^2 - 2 q.x, error bound
 * 1.02 (dp + 4) 2^-24 (|q| + |x|)^2, pairs inside it re-decided in FP64);
 * off (0): the direct-difference FP32 tiles; negative (the default): on for
 * rows of 32 dimensions or more (the bound is absolute in the norms, too wide
 * for low-dimensional rows in dense clusters).  Range calls with IDs always
 * use the direct tiles.  Results are identical either way. */
int gts_set_fp32_mfma(gts_index *idx, int on);

/* Memory budget of later query calls (range, kNN, update searches), in
 * bytes; 0 (the default) takes half of the free device memory at each call,
 * as searchIndexRnnV2 / searchIndexKnnV2 size their scratch
 * (search_v2.cuh:791-805, 984-1001).  A call whose search would exceed it
 * runs its query batch in sub-batches, split depth-first by query range
 * (search_v2.cuh:818-857, 921-966, 1177-1221): the per-query buffers (about
 * 32 KB a kNN query, at least 256 of them; 40 B a range query) and each
 * sub-batch's frontiers, leaf records, bands and range results stay within
 * the budget.  Index-resident data (rows, tree, cell partition, network) and
 * the caller's buffers are outside it.  Results do not depend on the budget;
 * a budget that cannot hold one query's search fails the call with
 * GTS_E_OOM.  bytes < 0 fails with GTS_E_ARG. */
int gts_set_memory_budget(gts_index *idx, int64_t bytes);

/* Rebuild cadence of later gts_update calls (no counterpart in the reference,
 * which rebuilds only when its cache list is full, update.cuh:484-540): with
 * nops > 0 the index is also rebuilt after every nops-th op of a stream (a
 * streaming workload's "rebuild every N ops"); 0 (the default) turns it off.
 * A rebuild keeps the logical order (live tree rows, then the cache list), so
 * counts and logical IDs do not depend on the cadence; physical row numbers
 * (insert and query op IDs) are those of the current tree. */
int gts_set_rebuild_period(gts_index *idx, int64_t nops);

/* Radius source of later learned kNN calls (learned = 1; no counterpart in
 * the reference): GTS_RADIUS_NETWORK (the default) predicts each query's
 * k-th NN distance with the trained network; GTS_RADIUS_PIVOTS takes the k-th
 * smallest FP32 distance to the pivots of one tree level (the deepest with at
 * most 4096 split nodes), rounded up -- the plain kNN bound of updateDisK
 * (search_v2.cuh:755-777) over a flat pivot level -- and needs no training.
 * A query whose k exceeds that pivot count gets an unbounded radius.  Results
 * are identical for either source. */
int gts_set_radius_source(gts_index *idx, int source);

/* Safety factor of the learned radius (no counterpart in the reference): the
 * predicted radius is scaled to cover quantile `q` (0 < q <= 1; 1 = the
 * largest) of the training queries' true / predicted radius ratios; the
 * default is 0.999.  Queries whose radius still falls short are rerun with a
 * doubled radius, so results are identical for every q; a higher q trades
 * reruns for more candidates. */
int gts_set_radius_quantile(gts_index *idx, double q);

/* Cell count of the cell scan (no counterpart in the reference): 0 (the
 * default) picks min(4096, n / 256); the partition is rebuilt at the next
 * learned kNN call that uses it.  Results are identical for every count. */
int gts_set_cell_count(gts_index *idx, int ncell);

/* Schedule autotuner of the dense verification kernel (no counterpart in the
 * reference): with `on` (the default) the first launch of each shape times
 * the candidate schedules (row-range size; one block per work item or one
 * persistent block per CU) on the batch itself and keeps the fastest; off uses
 * the default schedule.  Results are identical either way. */
int gts_set_autotune(gts_index *idx, int on);

/* Fixes the dense kernel's schedule for later calls (overrides the
 * autotuner): `rows` per row range (a multiple of 32; 0 returns to automatic
 * choice) and `persistent` = 1 for one block per CU.  A schedule whose LDS
 * tile does not fit fails the query with GTS_E_CAPACITY. */
int gts_set_dense_schedule(gts_index *idx, int persistent, int rows);

/* Graph replay of learned kNN calls (no counterpart in the reference; on by
 * default): once a batch shape has an exhaustive plan and a tuned dense
 * schedule, the second identical call (same sizes, buffers and parameters)
 * captures its first pass -- query preparation, radius network, dense
 * verification and top-k -- into a hipGraph and later identical calls replay
 * it.  Radius-doubling reruns run as ordinary launches.  Results are identical
 * either way. */
int gts_set_graph(gts_index *idx, int on);

/* Event timing of query calls (no counterpart in the reference; on by
 * default): ms_verify of gts_last_stats comes from HIP events recorded around
 * the verification kernels, and a kNN call's ms_total from events around the
 * call.  Off: neither is recorded (in a replayed kNN pass each event holds the
 * next kernel back for a few microseconds) and both report 0 (a range call
 * keeps its ms_total).  Results are identical either way. */
int gts_set_timing(gts_index *idx, int on);

/* The same with host buffers: `ks` may be NULL (then every query uses kfix). */
int gts_knn_query_ex(gts_index *idx, const int32_t *qids, int64_t nq, const int32_t *ks, int kfix,
                     int kstride, int32_t *ids, float *dists, float *kth, int learned, int verify);

/* Device buffers on the index's GPU for the *_dev entry points (no
 * counterpart in the reference: it uses managed memory throughout). */
int gts_device_alloc(gts_index *idx, int64_t bytes, void **ptr);
int gts_device_free(gts_index *idx, void *ptr);
int gts_copy_to_device(gts_index *idx, void *dst, const void *src, int64_t bytes);
int gts_copy_to_host(gts_index *idx, void *dst, const void *src, int64_t bytes);
int gts_copy_device(gts_index *idx, void *dst, const void *src, int64_t bytes);

/* ---- learned radius (no counterpart in the reference) -------------------
 * Trains the 16-64-64-1 GELU network (BF16 weights, FP32 accumulation) on the
 * GPU to map a query's distances to its 16 nearest level pivots to log of its
 * k-th NN distance; labels come from exact plain kNN over `ntrain` training
 * query IDs.  Returns the final training loss in *loss. */
int gts_learned_train(gts_index *idx, const int32_t *train_qids, int64_t ntrain, int kmax,
                      int epochs, float lr, uint64_t seed, float *loss);

/* Loss (mean of (y - t)^2 / 2) and gradients of the radius network on the GPU
 * for packed FP32 weights W1[64x16] b1[64] W2[64x64] b2[64] W3[64] b3[1]
 * (rounded to BF16 for the forward pass), inputs X[nb x 16] and targets t[nb]:
 * the gradient check of the training kernels.  gts_mlp_num_params() gives the
 * packed length. */
int gts_mlp_num_params(void);
int gts_mlp_loss_grad(gts_index *idx, const float *X, const float *t, int64_t nb, const float *w,
                      float *grad, float *loss);

/* No counterpart in the reference.  The same loss and gradients with X[nb x
 * 16], t[nb], w, grad and loss (one float) in device memory on the index's
 * GPU, allocating nothing -- for callers that cannot allocate (code under
 * graph capture, tight serving or training loops, harnesses that forbid it).
 * The activations live in a fixed device array of the library, one per GPU
 * (about 3.2 MB, held until the process ends): nb is at most
 * GTS_MLP_NOALLOC_MAX_NB (larger fails with GTS_E_ARG).  Every call on a GPU
 * shares that array, so calls that overlap on one GPU race.  Runs on the
 * index stream and returns when the outputs are complete. */
#define GTS_MLP_NOALLOC_MAX_NB 2048
int gts_mlp_loss_grad_noalloc(gts_index *idx, const float *d_X, const float *d_t, int64_t nb,
                              const float *d_w, float *d_grad, float *d_loss);

/* statistics of the last query call on this index */
typedef struct {
    int64_t frontier_pairs;   /* (query, node) ring tests                   */
    int64_t leaf_records;     /* (query, parent, leaf-mask) records         */
    int64_t verified_pairs;   /* (query, object) distance evaluations       */
    int64_t candidates;       /* kNN candidates kept for selection          */
    int64_t band_pairs;       /* pairs re-decided in FP64                   */
    int64_t retries;          /* radius doublings / capacity reruns         */
    double ms_total;          /* device time of the call                    */
    double ms_verify;         /* device time of the verification kernels    */
    int64_t verify_launches;  /* verification kernel launches               */
    int64_t exhaustive_passes; /* kNN passes planned without tree pruning    */
    int64_t dense_rows;       /* schedule of the call's largest dense-kernel */
    int64_t dense_persistent; /* launch: rows per range, persistent or grid  */
    int64_t graph_replays;    /* first passes run as a replayed hipGraph     */
    int64_t cell_scan;        /* cells of the cell scan the call used, or 0  */
    int64_t live_chunks;      /* cell scan: (64-query block, 32-row chunk) pairs verified */
    /* gts_update calls (update.cuh:449-638's time_update_s/_u, count_update_s/_u): */
    double ms_update_search;  /* wall time of the query ops' searches          */
    double ms_update_apply;   /* wall time of the inserts, deletes, rebuilds   */
    int64_t update_queries;   /* query ops (flag 2)                            */
    int64_t update_changes;   /* insert and delete ops                         */
    double ms_cells_build;    /* wall time of building the cell scan's partition in this call (0: none) */
    int64_t budget_batches;   /* sub-batches the memory budget split the call into (1: none) */
    /* the first pass's verification launches alone (the batch; reruns excluded): */
    double ms_verify_batch;
    int64_t verify_launches_batch;
    int64_t live_chunks_batch; /* cell scan: their (query block, chunk) pairs */
} gts_stats;
int gts_last_stats(const gts_index *idx, gts_stats *out);

#ifdef __cplusplus
}
#endif
#endif
