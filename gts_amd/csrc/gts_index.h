// The index object behind the C ABI handles.
#pragma once

#include <cstdint>
#include <map>
#include <vector>

#include "gts_internal.h"

struct gts_dataset {
    int64_t n = 0;
    int dim = 0;      // vector dimension, or the string stride for metric 6
    int metric = 2;
    std::vector<float> vec;      // n * dim (metrics 0, 1, 2, 5)
    std::vector<uint8_t> str;    // n * stride (metric 6)
    std::vector<int32_t> len;    // n (metric 6)
    int stride = 0;
};

namespace gts {

// Schedule of the dense verification kernel: rows per row range (a block's
// work item) and grid (one block per item) or persistent (one block per CU).
struct DenseSched {
    bool persist = false;
    int rows = 0;
};

// A captured first pass and the dense-kernel schedule it launches.
struct GraphEntry {
    uint64_t key = 0;
    hipGraphExec_t exec = nullptr;
    DenseSched sched;
};

// Cell scan of the learned kNN form: a Voronoi partition of the rows over
// sampled pivot rows and the rows copied in (cell, distance to pivot) order.
struct CellIndex {
    bool built = false;
    int ncell = 0;                          // cells (one pivot row each)
    int want = 0;                           // requested cell count (0: automatic)
    int32_t *cpos = nullptr;                // tree position of each cell's pivot row
    int32_t *scan_pos = nullptr;            // scan row -> tree position
    int32_t *pos_scan = nullptr;            // tree position -> scan row
    int32_t *sid = nullptr;                 // scan row -> object ID
    float *xn2 = nullptr;                   // |x|^2 by scan row
    uint8_t *rows = nullptr;                // scan-order rows, per 32-row chunk and K step one 3 KB piece:
                                            // BF16 hi fragments of rows 0-15, of rows 16-31, FP8 residual
                                            // fragments (x - hi scaled by 2^lo_shift) of all 32 rows
    int lo_shift = 0;
    float *eps = nullptr;                   // per scan row: |x - hi - lo|_2 of the decoded residual
    float eps_max = 0.0f;
    int2 *cell_rows = nullptr;              // per cell: first and last scan row (x = -1: empty)
    float *sdist = nullptr;                 // per scan row: distance to its cell's pivot row (FP32)
    uint16_t *phi = nullptr;                // BF16 hi pivot rows (fragment-major, cell order; the bounds use hi x hi)
    float *pn2 = nullptr;                   // |pivot|^2 by cell
    float4 *cell_pack = nullptr;            // per cell: |pivot|^2, |pivot|, conservative row distance range
};

// Learned-radius network: 16 -> 64 -> 64 -> 1, GELU, BF16 weights (kept as
// their FP32 master copies for training), FP32 accumulation.
struct LearnedModel {
    bool trained = false;
    int din = 16, dh = 64;
    int level = -1;          // tree level whose pivots give the features
    float scale = 1.0f;      // safety factor applied to the predicted radius
    float feat_mu = 0.f, feat_sd = 1.f, y_mu = 0.f, y_sd = 1.f;
    // device: FP32 master weights and their BF16 copies (inference)
    float *w = nullptr;      // W1[dh*din] b1[dh] W2[dh*dh] b2[dh] W3[dh] b3[1]
    uint16_t *wb = nullptr;  // BF16 copy of w
    int64_t nw = 0;
    int32_t *piv = nullptr;  // feature pivots (tree positions)
    int npiv = 0;
    // L2: BF16 hi copies of the feature pivot rows (fragment-major, padded
    // to 32 rows) and their norms, for the MFMA distance table
    uint16_t *phi = nullptr;
    float *pn2 = nullptr;
    float *mu = nullptr, *sd = nullptr;  // feature standardisation (device)
    int kmax = 0;            // the network predicts the kmax-th NN distance
    std::vector<float> resid;   // sorted log(true / predicted) radii of the training queries
    float quantile = 0.999f;    // the safety factor covers this quantile of them
};

}  // namespace gts

struct gts_index {
    int device = 0;
    hipStream_t st = nullptr;
    int64_t n = 0;
    int dim = 0, dp = 0, metric = 2;
    int T = 10, max_size = 20, max_h = 0, height = 0;
    int64_t num_nodes = 0;
    std::vector<int64_t> level_start;  // height + 1
    std::vector<int64_t> pnum;         // pnum[l] = split nodes at level l - 1
    gts::ErrModel em{};

    // device arrays
    float *X = nullptr;        // n x dp, tree order
    float *xn2 = nullptr;      // |x|^2 per position (FP32)
    uint16_t *Xhi = nullptr;   // BF16 high parts, n x dk (MFMA form)
    uint16_t *Xlo = nullptr;   // BF16 low parts
    int dk = 0;                // dp rounded up to the MFMA K step
    float xn2_max = 0.f;       // largest |x|^2 (error bound of the norm-expansion forms)
    bool xn2_max_ok = false;   // xn2_max holds for this tree
    int fp32_mfma = -1;        // FP32-form L2 tiles on the FP32 matrix cores (gts_set_fp32_mfma): 1 on, 0 off,
                               // -1 from FP32_MFMA_MIN_DIM dimensions on
    int64_t rebuild_period = 0;   // update streams: also rebuild after every period-th op (0: only when the
                                  // cache list is full)
    uint8_t *S = nullptr;
    int32_t *slen = nullptr;
    int sstride = 0;
    int32_t *perm = nullptr, *inv = nullptr;
    int32_t *nlid = nullptr, *nsize = nullptr, *npiv = nullptr;
    float *nmin = nullptr;
    uint8_t *nflag = nullptr;
    float *ring = nullptr;       // per node: children's min_dis, +inf sentinel
    uint32_t *cne = nullptr, *cleaf = nullptr;

    // host copy of the current physical rows (the update path rebuilds from it)
    gts_dataset data;

    // scratch reused across calls
    gts::DBuf<int32_t> fr_q[2], fr_n[2];       // frontier (query index, node)
    gts::DBuf<float> fr_dq;                      // d(q, pivot) per frontier pair
    gts::DBuf<int32_t> cnt_int, qcnt;           // children counts
    gts::DBuf<uint32_t> leafmask;
    gts::DBuf<int64_t> off_int, off_leaf, qoff;  // scans
    gts::DBuf<int32_t> rec_q, rec_n;            // leaf records
    gts::DBuf<uint32_t> rec_m;
    gts::DBuf<int32_t> grp_q, grp_n;            // records grouped by node
    gts::DBuf<uint32_t> grp_m;
    gts::DBuf<uint64_t> rs_k0, rs_k1;            // record sort by (query, parent): keys
    gts::DBuf<int32_t> rs_v0, rs_v1;             // and record indexes
    gts::DBuf<int32_t> node_cnt;
    gts::DBuf<int64_t> node_off, tile_off, tmp64;
    gts::DBuf<int32_t> tile_slot;                // FP32 MFMA tiles: slot of each tile
    gts::DBuf<int32_t> qpos;
    gts::DBuf<float> bnd;                        // per-query bound (r or kNN bound)
    gts::DBuf<int32_t> kq;                       // per-query k
    gts::DBuf<uint64_t> cand;                    // kNN candidates (dist bits << 32 | id)
    gts::DBuf<int32_t> ccount;
    gts::DBuf<int32_t> band_q, band_p;           // range band pairs
    gts::DBuf<int64_t> counts;                   // range counts
    gts::DBuf<uint64_t> res_k, res_k2;           // range results (q << 32 | id)
    gts::DBuf<float> res_d;
    gts::DBuf<int32_t> res_v, res_v2;
    gts::DBuf<int32_t> status;
    gts::DBuf<int64_t> scal;                     // small device scalars
    gts::DBuf<int32_t> dqid, dks, dids;
    gts::DBuf<float> ddist, dkth, dfeat, dpred;
    gts::DBuf<int32_t> orig;                     // output row of each batch-local query
    gts::DBuf<int32_t> fail;                     // kNN: queries of a pass that need a rerun
    int64_t *hpin = nullptr;                     // pinned coherent host block: per-pass scalars, [15] sequence
    int64_t pub_seq = 0;                         // last publish counter value seen in hpin[15]
    const int64_t *pub_scal = nullptr;           // scal block whose counter was last cleared
    bool pub_clean = true;                       // the last kNN call returned normally
    int pend_timing = 0;                         // event pairs of the last call read when stats are asked for
                                                 // (1: graph_ev verify, 2: timing_ev verify, 4: call_ev total)
    hipEvent_t call_ev[2] = {nullptr, nullptr};    // bracket a kNN call
    hipEvent_t timing_ev[2] = {nullptr, nullptr};  // bracket an ordinary pass's verification kernel
    bool timing_pending = false;                   // timing_ev recorded, not read yet
    bool timing = true;                            // kNN calls record their timing events
    gts::DBuf<float> dense;                      // dense query x pivot table / leaf-mask table
    gts::DBuf<int32_t> par_of;                   // parent node of each position at par_level
    gts::DBuf<int32_t> plan_q, plan_k;           // query sample of the kNN planner
    gts::DBuf<uint16_t> qimg;                    // BF16 LDS images of the dense kernel's query blocks
    gts::DBuf<float> ftab;                       // query x feature-pivot distance table
    gts::DBuf<float> plan_b;
    int64_t par_level = -1;                      // first node ID of the level par_of describes
    gts::DBuf<int32_t> death;                    // update path: deleting op index by position
    const int32_t *death_ptr = nullptr;          // deletions seen by the next search
    const int32_t *qop_ptr = nullptr;            // op index of each query of the next search
    int verify_form = GTS_VERIFY_FP32;
    int knn_plan = GTS_PLAN_AUTO;
    // automatic plan decisions by batch shape: (exhaustive?, batches since sampled)
    std::map<uint64_t, std::pair<bool, int>> plan_memo;
    // captured first passes of learned exhaustive kNN calls, by call key
    // (buffers, sizes, schedule), most recently used first (see knn_run)
    bool use_graph = true;
    std::vector<gts::GraphEntry> graphs;
    std::vector<uint64_t> graph_pending;             // keys seen once, captured when seen again
    hipEvent_t graph_ev[2] = {nullptr, nullptr};   // bracket the dense kernel of every graph
    bool autotune = true;                                  // time dense-kernel schedules on first use
    gts::DenseSched forced_sched;                          // rows > 0: schedule fixed by the caller
    gts::DenseSched last_sched;                            // schedule of the call's largest dense launch
    int64_t last_sched_nq = 0;
    std::map<uint64_t, gts::DenseSched> dense_sched;       // chosen schedule by launch shape
    std::map<uint64_t, int> lowdim_sched;                  // persistent low-dimensional pass: blocks per CU by shape
    gts::DBuf<uint32_t> ld_work;                           // its work counter

    gts::CellIndex cells;
    bool cells_plan = false;                     // the current kNN call verifies through the cell scan
    gts::DBuf<uint32_t> cell_bits;               // live-chunk bitmap rows of the query blocks
    gts::DBuf<int32_t> cell_list, cell_cnt, qperm, qrank, qtmp;
    gts::DBuf<int32_t> cell_pcnt;                // verification kernel's per-query counters (128 B apart)
    gts::DBuf<int32_t> argerr;                   // kNN argument check flag (published with each pass)
    gts::DBuf<uint64_t> ex_k;                    // exact kNN fallback: keys and IDs of every object (x 2)
    gts::DBuf<int32_t> ex_v;
    // radius source of learned kNN calls (GTS_RADIUS_*) and, for the pivot
    // source, its level pivots (rp_n < 0: not chosen yet for this tree)
    int radius_src = GTS_RADIUS_NETWORK;
    gts::DBuf<int32_t> rp_piv;
    gts::DBuf<int64_t> rp_off;
    int64_t rp_n = -1;
    int rp_level = -1;
    // memory budget of query calls (gts_set_memory_budget; 0: half of the
    // free device memory at each call) and, while a call runs, its limit and
    // the bytes of its current sub-batch's per-query buffers (limit 0: none)
    int64_t mem_budget = 0;
    int64_t budget_limit = 0, budget_fixed = 0;
    const int32_t *qrank_zero = nullptr;         // qrank block known to be all zero (its capacity below)
    size_t qrank_zero_cap = 0;
    gts::DBuf<float> qtmp_f;
    // the first pass graph's speculative rerun batch: qpos, kq, orig, status,
    // ccount (64 slots each), radii, candidates
    gts::DBuf<int32_t> rq_i;
    gts::DBuf<float> rq_b;
    gts::DBuf<uint64_t> rq_cand;

    gts_stats stats{};
    gts::LearnedModel model;

    gts::DevTree dev() const
    {
        gts::DevTree t;
        t.X = X; t.S = S; t.slen = slen; t.perm = perm; t.nlid = nlid; t.nsize = nsize;
        t.nmin = nmin; t.npiv = npiv; t.nflag = nflag; t.n = n; t.dim = dim; t.dp = dp;
        t.metric = metric; t.T = T; t.sstride = sstride; t.death = death_ptr; t.qop = qop_ptr;
        t.ring = ring; t.cne = cne; t.cleaf = cleaf; t.xn2 = xn2;
        return t;
    }
};

namespace gts {
// builds from src when given (idx->data may be filled meanwhile), else from idx->data
void build_tree(gts_index *idx, const gts_dataset *src = nullptr);   // gts_build.hip
void release_tree(gts_index *idx);
void graph_reset(gts_index *idx);         // gts_query.hip: drop the captured first pass
int tree_height(int64_t n, int T, int max_size, int max_h);

// Bytes of the running call's memory budget left for its sub-batch's search
// scratch (frontiers, leaf records, bands, range results), or INT64_MAX.
inline int64_t budget_room(const gts_index *I)
{
    return I->budget_limit > 0 ? I->budget_limit - I->budget_fixed : INT64_MAX;
}
// A working set of `bytes` must fit the budget, else the sub-batch is split.
inline void budget_check(const gts_index *I, int64_t bytes)
{
    if (I->budget_limit > 0 && bytes > budget_room(I)) throw BudgetExceeded{};
}
}  // namespace gts
