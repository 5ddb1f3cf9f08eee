/*
 * CPU restatement of the GTS pivot-tree algorithm -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is the oracle.  Only the test suite, the smoke check and the
 * CPU-baseline leg of the benchmark load it, always as the checker (or as the
 * timed CPU baseline), never as the product path.  The GPU library never links
 * or calls it.
 *
 * Everything is FP64 arithmetic on the FP32 inputs (edit distance is exact
 * integer DP), so result sets are defined by exact distances:
 *   range(q, r)  = { x : d(q, x) <= r }                (search_v2.cuh:416)
 *   knn(q, k)    = k smallest by (d, id)               (search_v2.cuh:1160-1169)
 * The query object itself is at distance 0 (search_v2.cuh:338-340).
 *
 * Restated from the reference (file:line under "Source Code/GTS/"):
 *   distances ............ include/tree.cuh:81-154, include/search_v2.cuh:341-414
 *   tree construction .... include/tree.cuh:53-393 (indexConstru)
 *   range two-stage ...... include/search_v2.cuh:40-151, 319-430, 780-972
 *   kNN two-stage ........ include/search_v2.cuh:155-196, 433-544, 600-777, 975-1228
 *   update cache list .... include/update.cuh:97-216, 449-638, search_naive.cuh:24-251
 *                          (logical IDs selected with a Fenwick tree; sampled
 *                          evaluation of query ops for long streams)
 *
 * Deliberate differences from the reference, shared with the GPU library:
 *   - sort ties inside a node are broken by object ID (the reference's thrust
 *     sort on a composite double key leaves tie order to the sort);
 *   - a node at the last allowed level is always a leaf (tree.cuh:282-291 marks
 *     oversized ones non-leaf and search_v2.cuh:332 then skips them), and so is
 *     a node with fewer than T objects (its split would repeat itself: T-1 empty
 *     children and one child of the same size);
 *   - the kNN bound is kept in the distance precision (search_v2.cuh:771-774
 *     rounds it to float) and "+infinity" is used instead of INFI_DIS = 10000;
 *   - angular cosines are clamped to [-1, 1] (tree.cuh:122-125 clamps only > 1).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_LINF 0
#define ORC_L1 1
#define ORC_L2 2
#define ORC_ANG 5
#define ORC_EDIT 6

typedef struct {
    int64_t n;
    int dim;
    int metric;
    const float *vec;      /* n * dim, row-major (metrics 0,1,2,5) */
    const uint8_t *str;    /* n * stride (metric 6)                 */
    const int32_t *len;    /* n string lengths (metric 6)           */
    int stride;
} orc_data;

/* ------------------------------------------------------------- distances */

/* Levenshtein distance with unit costs (tree.cuh:128-154: full Wagner-Fischer DP). */
int orc_edit(const uint8_t *a, int la, const uint8_t *b, int lb)
{
    if (la == 0) return lb;
    if (lb == 0) return la;
    int row[1025];
    if (lb > 1024) return -1;
    for (int j = 0; j <= lb; j++) row[j] = j;
    for (int i = 1; i <= la; i++) {
        int diag = row[0];
        row[0] = i;
        for (int j = 1; j <= lb; j++) {
            int up = row[j];
            int cost = (a[i - 1] == b[j - 1]) ? 0 : 1;
            int v = up + 1;
            if (row[j - 1] + 1 < v) v = row[j - 1] + 1;
            if (diag + cost < v) v = diag + cost;
            row[j] = v;
            diag = up;
        }
    }
    return row[lb];
}

/* FP64 distance between two FP32 rows (tree.cuh:81-127). */
double orc_dist_vec(const float *a, const float *b, int dim, int metric)
{
    double s = 0.0;
    if (metric == ORC_L2) {
        for (int j = 0; j < dim; j++) {
            double d = (double)a[j] - (double)b[j];
            s += d * d;
        }
        return sqrt(s);
    } else if (metric == ORC_L1) {
        for (int j = 0; j < dim; j++) s += fabs((double)a[j] - (double)b[j]);
        return s;
    } else if (metric == ORC_LINF) {
        for (int j = 0; j < dim; j++) {
            double d = fabs((double)a[j] - (double)b[j]);
            if (d > s) s = d;
        }
        return s;
    } else if (metric == ORC_ANG) {
        double na = 0, nb = 0, dot = 0;
        for (int j = 0; j < dim; j++) {
            na += (double)a[j] * a[j];
            nb += (double)b[j] * b[j];
            dot += (double)a[j] * b[j];
        }
        double c = dot / (sqrt(na) * sqrt(nb));
        if (c > 1.0) c = 1.0;
        if (c < -1.0) c = -1.0;
        return fabs(acos(c) * 180.0 / 3.1415926);
    }
    return NAN;
}

/* Distance between objects i and j; d(i, i) = 0 by identity (search_v2.cuh:338). */
double orc_dist(const orc_data *D, int64_t i, int64_t j)
{
    if (i == j) return 0.0;
    if (D->metric == ORC_EDIT)
        return (double)orc_edit(D->str + (size_t)i * D->stride, D->len[i],
                                D->str + (size_t)j * D->stride, D->len[j]);
    return orc_dist_vec(D->vec + (size_t)i * D->dim, D->vec + (size_t)j * D->dim, D->dim,
                        D->metric);
}

/* ------------------------------------------------------------ brute force */

void orc_bf_range_count(const orc_data *D, const int32_t *qids, int nq, double r,
                        int64_t *counts, int nthreads)
{
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
    for (int qi = 0; qi < nq; qi++) {
        int64_t c = 0;
        for (int64_t x = 0; x < D->n; x++)
            if (orc_dist(D, qids[qi], x) <= r) c++;
        counts[qi] = c;
    }
}

/* ids/dists at offsets[qi] .. offsets[qi+1], ascending by ID. */
void orc_bf_range_ids(const orc_data *D, const int32_t *qids, int nq, double r,
                      const int64_t *offsets, int32_t *ids, double *dists, int nthreads)
{
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
    for (int qi = 0; qi < nq; qi++) {
        int64_t o = offsets[qi];
        for (int64_t x = 0; x < D->n; x++) {
            double d = orc_dist(D, qids[qi], x);
            if (d <= r) {
                ids[o] = (int32_t)x;
                dists[o] = d;
                o++;
            }
        }
    }
}

typedef struct {
    double d;
    int32_t id;
} orc_pair;

static int pair_less(orc_pair a, orc_pair b)
{
    return a.d < b.d || (a.d == b.d && a.id < b.id);
}

/* Bounded max-heap of the k smallest (d, id) pairs. */
static void heap_push(orc_pair *h, int *sz, int k, orc_pair p)
{
    if (*sz < k) {
        int i = (*sz)++;
        h[i] = p;
        while (i > 0) {
            int par = (i - 1) / 2;
            if (pair_less(h[par], h[i])) {
                orc_pair t = h[par]; h[par] = h[i]; h[i] = t;
                i = par;
            } else break;
        }
    } else if (k > 0 && pair_less(p, h[0])) {
        h[0] = p;
        int i = 0;
        for (;;) {
            int l = 2 * i + 1, r = l + 1, m = i;
            if (l < k && pair_less(h[m], h[l])) m = l;
            if (r < k && pair_less(h[m], h[r])) m = r;
            if (m == i) break;
            orc_pair t = h[m]; h[m] = h[i]; h[i] = t;
            i = m;
        }
    }
}

static int pair_cmp(const void *a, const void *b)
{
    orc_pair x = *(const orc_pair *)a, y = *(const orc_pair *)b;
    if (pair_less(x, y)) return -1;
    if (pair_less(y, x)) return 1;
    return 0;
}

static void heap_emit(orc_pair *h, int sz, int k, int32_t *ids, double *dists)
{
    qsort(h, sz, sizeof(orc_pair), pair_cmp);
    for (int i = 0; i < k; i++) {
        ids[i] = i < sz ? h[i].id : -1;
        dists[i] = i < sz ? h[i].d : INFINITY;
    }
}

/* kNN by brute force: ids/dists [nq, k] ascending by (d, id). */
void orc_bf_knn(const orc_data *D, const int32_t *qids, int nq, int k, int32_t *ids,
                double *dists, int nthreads)
{
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
    {
        orc_pair *h = (orc_pair *)malloc(sizeof(orc_pair) * (size_t)(k > 0 ? k : 1));
#pragma omp for schedule(dynamic, 1)
        for (int qi = 0; qi < nq; qi++) {
            int sz = 0;
            for (int64_t x = 0; x < D->n; x++) {
                orc_pair p = {orc_dist(D, qids[qi], x), (int32_t)x};
                heap_push(h, &sz, k, p);
            }
            heap_emit(h, sz, k, ids + (size_t)qi * k, dists + (size_t)qi * k);
        }
        free(h);
    }
}

/* ---------------------------------------------------------- tree (a1-a6) */

typedef struct {
    int order, max_size, max_h, height;
    int64_t n, num_nodes;
    int64_t *level_start;  /* height + 1 entries: BFS index of first node per level */
    int32_t *perm;         /* position -> object ID                              */
    int32_t *lid, *size;   /* node range [lid, lid + size)                       */
    int32_t *pivot;        /* pivot object of a split node, -1 otherwise         */
    int32_t *ppivot;       /* pivot of the parent (TN.pid, tree.cuh:198)         */
    double *min_dis;       /* d(parent pivot, first object) (tree.cuh:205-280)   */
    int8_t *exists, *is_leaf;
} orc_tree;

/* Height of the count-based split for n objects (depends on sizes only). */
static int tree_height(int64_t n, int T, int max_size, int max_h)
{
    /* distinct node sizes per level; a level holds at most a handful */
    int64_t cur[64], nxt[64];
    int nc = 1, h = 1;
    cur[0] = n;
    for (;;) {
        if (max_h > 0 && h >= max_h) break;
        int nn = 0, any = 0;
        for (int i = 0; i < nc; i++) {
            int64_t s = cur[i];
            if (s <= max_size || s < T) continue;
            any = 1;
            int64_t avg = s / T, last = s - (T - 1) * avg;
            int64_t cand[2] = {avg, last};
            for (int c = 0; c < 2; c++) {
                int dup = 0;
                for (int j = 0; j < nn; j++) dup |= (nxt[j] == cand[c]);
                if (!dup && nn < 64) nxt[nn++] = cand[c];
            }
        }
        if (!any) break;
        memcpy(cur, nxt, sizeof(int64_t) * nn);
        nc = nn;
        h++;
    }
    return h;
}

typedef struct {
    double key;
    int32_t id;
} orc_kv;

static int kv_cmp(const void *a, const void *b)
{
    const orc_kv *x = (const orc_kv *)a, *y = (const orc_kv *)b;
    if (x->key < y->key) return -1;
    if (x->key > y->key) return 1;
    return (x->id > y->id) - (x->id < y->id);
}

void orc_tree_free(orc_tree *t)
{
    free(t->level_start); free(t->perm); free(t->lid); free(t->size); free(t->pivot);
    free(t->ppivot); free(t->min_dis); free(t->exists); free(t->is_leaf);
    memset(t, 0, sizeof(*t));
}

/* indexConstru (tree.cuh:332-393): level-by-level split of every oversized node.
 * Pivot = object at the middle position of the node's range (tree.cuh:73); the
 * range is sorted by (d(x, pivot), id) and cut into T equal-count children, the
 * last one taking the remainder (tree.cuh:193-202).  Returns 0 or -1 on bad args. */
int orc_tree_build(const orc_data *D, int T, int max_size, int max_h, orc_tree *t, int nthreads)
{
    memset(t, 0, sizeof(*t));
    if (T < 2 || max_size < 1 || D->n < 1 || max_h < 0) return -1;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    int h = tree_height(D->n, T, max_size, max_h);
    int64_t nodes = 0, pw = 1;
    t->level_start = (int64_t *)malloc(sizeof(int64_t) * (h + 1));
    for (int l = 0; l < h; l++) {
        t->level_start[l] = nodes;
        nodes += pw;
        pw *= T;
    }
    t->level_start[h] = nodes;
    t->order = T; t->max_size = max_size; t->max_h = max_h; t->height = h;
    t->n = D->n; t->num_nodes = nodes;
    t->perm = (int32_t *)malloc(sizeof(int32_t) * D->n);
    t->lid = (int32_t *)calloc(nodes, sizeof(int32_t));
    t->size = (int32_t *)calloc(nodes, sizeof(int32_t));
    t->pivot = (int32_t *)malloc(sizeof(int32_t) * nodes);
    t->ppivot = (int32_t *)malloc(sizeof(int32_t) * nodes);
    t->min_dis = (double *)calloc(nodes, sizeof(double));
    t->exists = (int8_t *)calloc(nodes, 1);
    t->is_leaf = (int8_t *)calloc(nodes, 1);
    for (int64_t i = 0; i < nodes; i++) { t->pivot[i] = -1; t->ppivot[i] = -1; }
    for (int64_t i = 0; i < D->n; i++) t->perm[i] = (int32_t)i;
    t->exists[0] = 1;
    t->size[0] = (int32_t)D->n;
    t->is_leaf[0] = (D->n <= max_size) || D->n < T || h == 1;
    for (int l = 0; l + 1 < h; l++) {
        int64_t s0 = t->level_start[l], s1 = t->level_start[l + 1];
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t nd = s0; nd < s1; nd++) {
            if (!t->exists[nd] || t->is_leaf[nd]) continue;
            int32_t lid = t->lid[nd], size = t->size[nd];
            int32_t piv = t->perm[lid + (size - 1) / 2];
            t->pivot[nd] = piv;
            orc_kv *kv = (orc_kv *)malloc(sizeof(orc_kv) * size);
            for (int32_t i = 0; i < size; i++) {
                int32_t x = t->perm[lid + i];
                kv[i].id = x;
                kv[i].key = orc_dist(D, x, piv);
            }
            qsort(kv, size, sizeof(orc_kv), kv_cmp);
            for (int32_t i = 0; i < size; i++) t->perm[lid + i] = kv[i].id;
            int32_t avg = size / T;
            for (int c = 0; c < T; c++) {
                int64_t ch = nd * T + c + 1;
                int32_t clid = lid + avg * c;
                int32_t csz = (c < T - 1) ? avg : size - (T - 1) * avg;
                t->lid[ch] = clid;
                t->size[ch] = csz;
                t->ppivot[ch] = piv;
                t->min_dis[ch] = kv[clid - lid].key;
                t->exists[ch] = 1;
                t->is_leaf[ch] = (csz <= max_size) || csz < T || (l + 1 == h - 1);
            }
            free(kv);
        }
    }
    return 0;
}

/* ------------------------------------------------- tree search (a8 - a20) */

/* Ring lower bound of child c of parent p (search_v2.cuh:139-145). */
static double ring_lb(const orc_tree *t, int64_t c, double dq)
{
    int64_t first = (c - 1) / t->order * t->order + 1;
    int i = (int)(c - first);
    double lb = t->min_dis[c] - dq;
    if (lb < 0) lb = 0;
    if (i < t->order - 1) {
        double lb2 = dq - t->min_dis[c + 1];
        if (lb2 > lb) lb = lb2;
    }
    return lb;
}

/* The FP64 oracle's own pruning slack: FP64 rounding of d and of the ring keys. */
static double prune_slack(double a, double b) { return 1e-12 * (fabs(a) + fabs(b)) + 1e-300; }

/* Range search over the tree (searchIndexRnnV2 semantics).  Per query: count of
 * results, and the number of verified objects and (query, node) tests. */
void orc_tree_range(const orc_data *D, const orc_tree *t, const int32_t *qids, int nq, double r,
                    int64_t *counts, int64_t *verified, int64_t *tested, int nthreads)
{
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
    {
        int64_t *fr = (int64_t *)malloc(sizeof(int64_t) * t->num_nodes);
        int64_t *nx = (int64_t *)malloc(sizeof(int64_t) * t->num_nodes);
#pragma omp for schedule(dynamic, 1)
        for (int qi = 0; qi < nq; qi++) {
            int32_t q = qids[qi];
            int64_t cnt = 0, ver = 0, tst = 0, nf = 0;
            /* leaves at any level are verified as they are reached */
            fr[nf++] = 0;
            for (int l = 0; l < t->height && nf > 0; l++) {
                int64_t nn = 0;
                for (int64_t f = 0; f < nf; f++) {
                    int64_t nd = fr[f];
                    if (t->is_leaf[nd]) {
                        for (int32_t p = t->lid[nd]; p < t->lid[nd] + t->size[nd]; p++) {
                            ver++;
                            if (orc_dist(D, q, t->perm[p]) <= r) cnt++;
                        }
                        continue;
                    }
                    double dq = orc_dist(D, q, t->pivot[nd]);
                    for (int c = 0; c < t->order; c++) {
                        int64_t ch = nd * t->order + c + 1;
                        tst++;
                        double lb = ring_lb(t, ch, dq);
                        if (lb <= r + prune_slack(r, dq)) nx[nn++] = ch;
                    }
                }
                int64_t *tmp = fr; fr = nx; nx = tmp;
                nf = nn;
            }
            counts[qi] = cnt;
            if (verified) verified[qi] = ver;
            if (tested) tested[qi] = tst;
        }
        free(fr);
        free(nx);
    }
}

static int dbl_cmp(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

/* kNN search over the tree (searchIndexKnnV2 semantics, search_v2.cuh:975-1228):
 * the bound disk starts at +inf; at node level l (children level) it is
 * tightened to the k-th smallest d(q, parent pivot) over the query's surviving
 * parents once l >= 3 and the level has >= k parents (search_v2.cuh:1065-1108);
 * children are kept iff ring_lb <= disk (search_v2.cuh:192); leaf objects with
 * d <= disk are candidates (search_v2.cuh:532) and the k smallest by (d, id)
 * are returned. */
void orc_tree_knn(const orc_data *D, const orc_tree *t, const int32_t *qids, int nq, const int32_t *ks,
                  int kfix, int32_t *ids, double *dists, int kstride, int64_t *verified,
                  int nthreads)
{
    if (nthreads > 0) omp_set_num_threads(nthreads);
    /* parents per level: number of split nodes at level l - 1 */
    int64_t *pnum = (int64_t *)calloc(t->height + 1, sizeof(int64_t));
    for (int l = 1; l < t->height; l++)
        for (int64_t nd = t->level_start[l - 1]; nd < t->level_start[l]; nd++)
            if (t->exists[nd] && !t->is_leaf[nd]) pnum[l]++;
#pragma omp parallel
    {
        int64_t *fr = (int64_t *)malloc(sizeof(int64_t) * t->num_nodes);
        int64_t *nx = (int64_t *)malloc(sizeof(int64_t) * t->num_nodes);
        double *dq = (double *)malloc(sizeof(double) * t->num_nodes);
        double *tmp = (double *)malloc(sizeof(double) * t->num_nodes);
        int64_t *leaves = (int64_t *)malloc(sizeof(int64_t) * t->num_nodes);
        orc_pair *h = NULL;
        int hcap = 0;
#pragma omp for schedule(dynamic, 1)
        for (int qi = 0; qi < nq; qi++) {
            int k = ks ? ks[qi] : kfix;
            if (k > hcap) { free(h); hcap = k; h = (orc_pair *)malloc(sizeof(orc_pair) * hcap); }
            int32_t q = qids[qi];
            double disk = INFINITY;
            int active = 0;
            int64_t nf = 0, nl = 0, ver = 0;
            if (t->is_leaf[0]) leaves[nl++] = 0;
            else fr[nf++] = 0;
            for (int l = 1; l < t->height && nf > 0; l++) {
                for (int64_t f = 0; f < nf; f++) dq[f] = orc_dist(D, q, t->pivot[fr[f]]);
                if (!active && !(pnum[l] < k || l <= 2)) active = 1;
                if (active && pnum[l] >= k && nf >= k) {
                    memcpy(tmp, dq, sizeof(double) * nf);
                    qsort(tmp, nf, sizeof(double), dbl_cmp);
                    if (tmp[k - 1] < disk) disk = tmp[k - 1];
                }
                int64_t nn = 0;
                for (int64_t f = 0; f < nf; f++) {
                    for (int c = 0; c < t->order; c++) {
                        int64_t ch = fr[f] * t->order + c + 1;
                        double lb = ring_lb(t, ch, dq[f]);
                        if (lb <= disk + prune_slack(disk, dq[f])) {
                            if (t->is_leaf[ch]) leaves[nl++] = ch;
                            else nx[nn++] = ch;
                        }
                    }
                }
                int64_t *sw = fr; fr = nx; nx = sw;
                nf = nn;
            }
            int sz = 0;
            for (int64_t li = 0; li < nl; li++) {
                int64_t nd = leaves[li];
                for (int32_t p = t->lid[nd]; p < t->lid[nd] + t->size[nd]; p++) {
                    ver++;
                    double d = orc_dist(D, q, t->perm[p]);
                    if (d <= disk + prune_slack(disk, d)) {
                        orc_pair pr = {d, t->perm[p]};
                        heap_push(h, &sz, k, pr);
                    }
                }
            }
            heap_emit(h, sz, k, ids + (size_t)qi * kstride, dists + (size_t)qi * kstride);
            if (verified) verified[qi] = ver;
        }
        free(fr); free(nx); free(dq); free(tmp); free(leaves); free(h);
    }
    free(pnum);
}

/* ------------------------------------------------- update path (a21 - a27) */

/* updateIndexRnn (update.cuh:449-638) as a state machine over physical rows.
 *   op 0 insert: append physical row `id` of the current data to the cache;
 *                when the cache holds `cache_cap` entries, the live tree rows
 *                (in physical order) followed by the cached rows become the new
 *                data and the tree is rebuilt (update.cuh:484-540).
 *   op 1 delete: `id` is a logical ID = rank among live tree rows, then cache
 *                entries (update.cuh:118-143); a tree row is tombstoned, a cache
 *                entry is removed (update.cuh:546-582).
 *   op 2 query : `id` is a physical row of the current data (update.cuh:588);
 *                the count is #live tree rows within r plus #cache entries
 *                within r (update.cuh:590-609).
 * `rows` holds the initial n0 rows; `fresh` (optional, metric != 6) holds rows
 * that insert ops with flag 3 append by index into `fresh` (an extension for
 * streams of new vectors).  counts_out receives one count per query op;
 * returns the number of query ops or -(index+1) of the first invalid op. */
/* Fenwick tree over live flags: rank -> index selection for logical IDs
 * (findIdx, update.cuh:118-143, counts live rows in physical order). */
typedef struct {
    int32_t *t;
    int64_t n;
} orc_fenwick;

static void fw_init(orc_fenwick *f, int64_t n)
{
    f->n = n;
    f->t = (int32_t *)calloc(n + 1, sizeof(int32_t));
    for (int64_t i = 1; i <= n; i++) {
        f->t[i] += 1;
        int64_t j = i + (i & -i);
        if (j <= n) f->t[j] += f->t[i];
    }
}
static void fw_add(orc_fenwick *f, int64_t i, int v)
{
    for (i++; i <= f->n; i += i & -i) f->t[i] += v;
}
/* index of the live row of rank k (0-based) */
static int64_t fw_select(const orc_fenwick *f, int64_t k)
{
    int64_t pos = 0, step = 1;
    while (step * 2 <= f->n) step *= 2;
    for (; step > 0; step >>= 1)
        if (pos + step <= f->n && f->t[pos + step] <= k) {
            pos += step;
            k -= f->t[pos];
        }
    return pos;
}

/* updateIndexRnn restated (update.cuh:449-638), sequentially, op by op:
 *   flag 0 insert : physical row `id` of the current data joins the cache
 *                   list; when the list holds cache_cap entries the data is
 *                   rebuilt as the live tree rows in physical order followed
 *                   by the cache entries (update.cuh:481-540);
 *   flag 3 insert : row `id` of `fresh` joins the cache list (extension);
 *   flag 1 delete : `id` is logical: rank among live tree rows, then cache
 *                   entries (update.cuh:546-582);
 *   flag 2 query  : `id` is a physical row; the count of live tree rows and
 *                   cache entries within r, the row itself at distance 0.
 * Queries with eval[op] == 0 (eval non-null) are not evaluated: their count
 * is -1.  A rebuild also follows every period-th op when period > 0 (the
 * benchmark stream's cadence; a rebuild keeps the logical order, so only
 * physical row numbers change).  Returns the number of query ops, or
 * -(op + 1) for an invalid op. */
int64_t orc_update3(const orc_data *D0, const float *fresh, int64_t nfresh, const int32_t *flags,
                    const int32_t *oids, int64_t nops, double r, int cache_cap, int64_t period,
                    const int8_t *eval, int64_t *counts_out, int nthreads)
{
    if (nthreads > 0) omp_set_num_threads(nthreads);
    int dim = D0->dim;
    int is_str = D0->metric == ORC_EDIT;
    int64_t cap_rows = D0->n + nops + 1;
    /* physical storage: each row is a handle into the base objects (0..n0-1) or
     * into `fresh` (n0 + i) */
    int64_t *rows = (int64_t *)malloc(sizeof(int64_t) * cap_rows);
    int64_t *nrows = (int64_t *)malloc(sizeof(int64_t) * cap_rows);
    int8_t *dead = (int8_t *)calloc(cap_rows, 1);
    int64_t *cache = (int64_t *)malloc(sizeof(int64_t) * (cache_cap + 1));
    int64_t tree_size = D0->n, in_size = 0, nq_out = 0, live = D0->n;
    for (int64_t i = 0; i < D0->n; i++) rows[i] = i;
    orc_fenwick fw;
    fw_init(&fw, tree_size);

    for (int64_t op = 0; op < nops; op++) {
        int32_t fl = flags[op], id = oids[op];
        if (fl == 0 || fl == 3) {
            int64_t hnd;
            if (fl == 0) {
                if (id < 0 || id >= tree_size) { nq_out = -(op + 1); break; }
                hnd = rows[id];
            } else {
                if (is_str || !fresh || id < 0 || id >= nfresh) { nq_out = -(op + 1); break; }
                hnd = D0->n + id;
            }
            cache[in_size++] = hnd;
        } else if (fl == 1) {
            if (id < 0 || id >= live + in_size) { nq_out = -(op + 1); break; }
            if (id < live) {
                int64_t found = fw_select(&fw, id);
                dead[found] = 1;
                fw_add(&fw, found, -1);
                live--;
            } else {
                int64_t c = (int64_t)id - live;
                for (int64_t j = c; j + 1 < in_size; j++) cache[j] = cache[j + 1];
                in_size--;
            }
        } else if (fl == 2) {
            if (id < 0 || id >= tree_size) { nq_out = -(op + 1); break; }
            if (eval && !eval[op]) {
                counts_out[nq_out++] = -1;
                goto next;
            }
            int64_t qh = rows[id];
            int64_t cnt = 0;
#pragma omp parallel for reduction(+ : cnt) schedule(static)
            for (int64_t i = 0; i < tree_size + in_size; i++) {
                int64_t xh;
                if (i < tree_size) {
                    if (dead[i]) continue;
                    xh = rows[i];
                } else xh = cache[i - tree_size];
                double d;
                if (i == id) d = 0.0;   /* the query row itself (search_v2.cuh:338) */
                else if (is_str)
                    d = (double)orc_edit(D0->str + (size_t)qh * D0->stride, D0->len[qh],
                                         D0->str + (size_t)xh * D0->stride, D0->len[xh]);
                else {
                    const float *a = qh < D0->n ? D0->vec + (size_t)qh * dim : fresh + (size_t)(qh - D0->n) * dim;
                    const float *b = xh < D0->n ? D0->vec + (size_t)xh * dim : fresh + (size_t)(xh - D0->n) * dim;
                    d = orc_dist_vec(a, b, dim, D0->metric);
                }
                if (d <= r) cnt++;
            }
            counts_out[nq_out++] = cnt;
        } else {
            nq_out = -(op + 1);
            break;
        }
    next:
        /* rebuild when the cache list is full (update.cuh:484-540) or, with
         * a period, after every period-th op */
        if (in_size == cache_cap || (period > 0 && (op + 1) % period == 0)) {
            int64_t m = 0;
            for (int64_t i = 0; i < tree_size; i++)
                if (!dead[i]) nrows[m++] = rows[i];
            for (int64_t c = 0; c < in_size; c++) nrows[m++] = cache[c];
            int64_t *sw = rows; rows = nrows; nrows = sw;
            tree_size = m;
            live = m;
            memset(dead, 0, cap_rows);
            in_size = 0;
            free(fw.t);
            fw_init(&fw, tree_size);
        }
    }
    free(fw.t);
    free(rows); free(nrows); free(dead); free(cache);
    return nq_out;
}

int64_t orc_update2(const orc_data *D0, const float *fresh, int64_t nfresh, const int32_t *flags,
                    const int32_t *oids, int64_t nops, double r, int cache_cap, const int8_t *eval,
                    int64_t *counts_out, int nthreads)
{
    return orc_update3(D0, fresh, nfresh, flags, oids, nops, r, cache_cap, 0, eval, counts_out, nthreads);
}

int64_t orc_update(const orc_data *D0, const float *fresh, int64_t nfresh, const int32_t *flags,
                   const int32_t *oids, int64_t nops, double r, int cache_cap, int64_t *counts_out,
                   int nthreads)
{
    return orc_update3(D0, fresh, nfresh, flags, oids, nops, r, cache_cap, 0, NULL, counts_out, nthreads);
}

/* Turns a drawn op stream into concrete (flag, id) ops under the state rules
 * of orc_update2: kind 0 -> (3, i) inserts the i-th fresh row; kind 1 ->
 * (1, floor(u (live + cache))) deletes a uniform live logical ID; kind 2 ->
 * (2, physical row of the live tree row of rank floor(u live)) queries a
 * uniform live tree object.  A delete or query with nothing live becomes a
 * query of row 0 / is skipped (flag -1).  With period > 0 the state is
 * rebuilt after every period-th emitted op as well, as orc_update3 does over
 * the stream without its skipped ops.  Returns the number of fresh rows used. */
int64_t orc_update_stream2(int64_t n0, int64_t nfresh, const int8_t *kind, const double *u, int64_t nops,
                           int cache_cap, int64_t period, int32_t *flags, int32_t *ids)
{
    int64_t cap_rows = n0 + nops + 1;
    int8_t *dead = (int8_t *)calloc(cap_rows, 1);
    int64_t tree_size = n0, live = n0, in_size = 0, nins = 0, emitted = 0;
    orc_fenwick fw;
    fw_init(&fw, tree_size);
    for (int64_t op = 0; op < nops; op++) {
        flags[op] = -1;
        ids[op] = 0;
        if (kind[op] == 0) {
            if (nins >= nfresh) continue;
            flags[op] = 3;
            ids[op] = (int32_t)nins++;
            in_size++;
        } else if (kind[op] == 1) {
            int64_t tot = live + in_size;
            if (tot == 0) continue;
            int64_t id = (int64_t)(u[op] * (double)tot);
            if (id >= tot) id = tot - 1;
            flags[op] = 1;
            ids[op] = (int32_t)id;
            if (id < live) {
                int64_t f = fw_select(&fw, id);
                dead[f] = 1;
                fw_add(&fw, f, -1);
                live--;
            } else {
                in_size--;
            }
        } else {
            if (live == 0) continue;
            int64_t k = (int64_t)(u[op] * (double)live);
            if (k >= live) k = live - 1;
            flags[op] = 2;
            ids[op] = (int32_t)fw_select(&fw, k);
        }
        /* the replay's rebuild rule, counted over the emitted ops */
        emitted++;
        if (in_size == cache_cap || (period > 0 && emitted % period == 0)) {
            tree_size = live + in_size;
            live = tree_size;
            in_size = 0;
            memset(dead, 0, cap_rows);
            free(fw.t);
            fw_init(&fw, tree_size);
        }
    }
    free(fw.t);
    free(dead);
    return nins;
}

int64_t orc_update_stream(int64_t n0, int64_t nfresh, const int8_t *kind, const double *u, int64_t nops,
                          int cache_cap, int32_t *flags, int32_t *ids)
{
    return orc_update_stream2(n0, nfresh, kind, u, nops, cache_cap, 0, flags, ids);
}

/* ---------------------------------------------------- learned-radius MLP */

/* Reference forward/backward of the radius network in FP64 (test pin for the
 * GPU training kernels): y = W3 gelu(W2 gelu(W1 x + b1) + b2) + b3 with the
 * tanh-free exact GELU x * Phi(x).  Weights are given as FP32 (the BF16 weights
 * widened).  loss = mean (y - t)^2 / 2.  Gradients are written to g* (same
 * shapes as the weights).  Returns the loss. */
static double gelu(double x) { return 0.5 * x * (1.0 + erf(x / sqrt(2.0))); }
static double gelu_d(double x)
{
    return 0.5 * (1.0 + erf(x / sqrt(2.0))) + x * exp(-0.5 * x * x) / sqrt(2.0 * M_PI);
}

double orc_mlp_loss_grad(const float *X, const float *tgt, int64_t nb, int din, int dh,
                         const float *W1, const float *b1, const float *W2, const float *b2,
                         const float *W3, const float *b3, double *gW1, double *gb1,
                         double *gW2, double *gb2, double *gW3, double *gb3, double *yout)
{
    memset(gW1, 0, sizeof(double) * dh * din);
    memset(gb1, 0, sizeof(double) * dh);
    memset(gW2, 0, sizeof(double) * dh * dh);
    memset(gb2, 0, sizeof(double) * dh);
    memset(gW3, 0, sizeof(double) * dh);
    gb3[0] = 0;
    double *z1 = (double *)malloc(sizeof(double) * dh), *a1 = (double *)malloc(sizeof(double) * dh);
    double *z2 = (double *)malloc(sizeof(double) * dh), *a2 = (double *)malloc(sizeof(double) * dh);
    double *d2 = (double *)malloc(sizeof(double) * dh), *d1 = (double *)malloc(sizeof(double) * dh);
    double loss = 0;
    for (int64_t s = 0; s < nb; s++) {
        const float *x = X + s * din;
        for (int i = 0; i < dh; i++) {
            double z = b1[i];
            for (int j = 0; j < din; j++) z += (double)W1[i * din + j] * x[j];
            z1[i] = z;
            a1[i] = gelu(z);
        }
        for (int i = 0; i < dh; i++) {
            double z = b2[i];
            for (int j = 0; j < dh; j++) z += (double)W2[i * dh + j] * a1[j];
            z2[i] = z;
            a2[i] = gelu(z);
        }
        double y = b3[0];
        for (int j = 0; j < dh; j++) y += (double)W3[j] * a2[j];
        if (yout) yout[s] = y;
        double e = y - tgt[s];
        loss += 0.5 * e * e;
        double gy = e / nb;
        gb3[0] += gy;
        for (int j = 0; j < dh; j++) {
            gW3[j] += gy * a2[j];
            d2[j] = gy * W3[j] * gelu_d(z2[j]);
        }
        for (int i = 0; i < dh; i++) {
            gb2[i] += d2[i];
            for (int j = 0; j < dh; j++) gW2[i * dh + j] += d2[i] * a1[j];
        }
        for (int j = 0; j < dh; j++) {
            double acc = 0;
            for (int i = 0; i < dh; i++) acc += d2[i] * W2[i * dh + j];
            d1[j] = acc * gelu_d(z1[j]);
        }
        for (int i = 0; i < dh; i++) {
            gb1[i] += d1[i];
            for (int j = 0; j < din; j++) gW1[i * din + j] += d1[i] * x[j];
        }
    }
    free(z1); free(a1); free(z2); free(a2); free(d2); free(d1);
    return loss / nb;
}
