// Pivot-tree construction on gfx950 (replaces indexConstru, tree.cuh:332-393).
//
// Level by level, every splitting node picks the object at the middle of its
// position range as pivot (tree.cuh:73), every object of a splitting node gets
// its distance to that pivot (getPivotDis, tree.cuh:53-171), each node's range
// is sorted by (distance, object ID), and the node is cut into T equal-count
// children whose min_dis is the distance of their first object (nodeSplit,
// tree.cuh:173-297).
//
// The reference emulates the per-node sort with one global thrust sort on the
// double key d/INFI_DIS + node*DIS_CODE (tree.cuh:155, 372).  Here each level
// computes the FP32 distance keys in position order (four objects per wave,
// 16-byte loads) and sorts every splitting node's range by (distance, object
// ID):
//  * nodes of at most LDS_SORT_MAX objects: one workgroup per node sorts the
//    (distance bits << 32 | ID) words of its range in LDS (bitonic network),
//    so the level needs no global sort pass;
//  * levels with a larger node: one stable LSD radix sort over all objects in
//    ID order (ties fall in ID order) on (segment << 32 | distance bits),
//    where the segment is the node's rank at its level when every object
//    lies in a splitting node of the level (a few bits: four passes at the
//    root) and otherwise the start of the object's range (finished leaves
//    keyed by offset, so they keep their order).
#include "gts_cells.h"
#include "gts_dist.h"
#include "gts_index.h"

namespace gts {

int tree_height(int64_t n, int T, int max_size, int max_h)
{
    std::vector<int64_t> cur{n}, nxt;
    int h = 1;
    for (;;) {
        if (max_h > 0 && h >= max_h) break;
        nxt.clear();
        bool any = false;
        for (int64_t s : cur) {
            if (s <= max_size || s < T) continue;
            any = true;
            int64_t avg = s / T, last = s - (int64_t)(T - 1) * avg;
            for (int64_t c : {avg, last}) {
                bool dup = false;
                for (int64_t v : nxt) dup |= (v == c);
                if (!dup) nxt.push_back(c);
            }
        }
        if (!any) break;
        cur.swap(nxt);
        h++;
    }
    return h;
}

// Per level of the tree (as tree_height simulates it): the largest size of
// a node that splits there (0: none), and whether every object lies in a
// splitting node of that level (no leaf at it or above).
static void level_plan(int64_t n, int T, int max_size, int h, std::vector<int64_t> &maxsplit,
                       std::vector<char> &allsplit)
{
    maxsplit.assign(h, 0);
    allsplit.assign(h, 0);
    std::vector<int64_t> cur{n}, nxt;
    bool all = true;
    for (int l = 0; l + 1 < h; l++) {
        nxt.clear();
        for (int64_t sz : cur) {
            if (sz <= max_size || sz < T) {
                all = false;
                continue;
            }
            maxsplit[l] = std::max(maxsplit[l], sz);
            const int64_t avg = sz / T, last = sz - (int64_t)(T - 1) * avg;
            for (int64_t c : {avg, last})
                if (std::find(nxt.begin(), nxt.end(), c) == nxt.end()) nxt.push_back(c);
        }
        allsplit[l] = all ? 1 : 0;
        cur.swap(nxt);
    }
}

namespace {

constexpr int64_t LDS_SORT_MAX = 16384;   // largest node sorted by one workgroup in LDS (128 KB)

unsigned grid_for(int64_t n, int block, int64_t cap = 1 << 20)
{
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

__device__ __forceinline__ bool is_split(uint8_t f) { return (f & 3) == 1; }

__global__ void k_pivots(const int32_t *perm, const int32_t *nlid, const int32_t *nsize,
                         const uint8_t *nflag, int32_t *npiv_id, int64_t s0, int64_t s1)
{
    int64_t nd = s0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (nd >= s1) return;
    if (is_split(nflag[nd])) npiv_id[nd] = perm[nlid[nd] + (nsize[nd] - 1) / 2];
}

// Distance keys in position order, 16 lanes per position (four per wave),
// for rows of dim floats with dim % 4 == 0 (16-byte aligned rows): lane gl
// takes the float4 groups gl, gl + 16, ... of both rows, five groups of each
// in flight at a time, then the 16 lanes combine their partial results.
// Positions outside the level's splitting nodes are skipped (their keys are
// not read).
template <int M>
__global__ void __launch_bounds__(256) k_keys_pos4(const float *__restrict__ X0, int dim, const int32_t *perm,
                                                   const int32_t *pos_node, const uint8_t *nflag,
                                                   const int32_t *npiv_id, int64_t s0, int64_t s1, int64_t n,
                                                   uint32_t *dkey)
{
    constexpr int G = 16, B = 5;
    const int gl = threadIdx.x % G;
    const int n4 = dim >> 2;
    const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    const int64_t ng = ((int64_t)gridDim.x * blockDim.x) / G;
    for (int64_t pos = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G; pos < n; pos += ng) {
        const int32_t nd = pos_node[pos];
        if (!(nd >= s0 && nd < s1 && is_split(nflag[nd]))) continue;   // the whole group
        const float4 *a = (const float4 *)(X0 + (int64_t)perm[pos] * dim);
        const float4 *b = (const float4 *)(X0 + (int64_t)npiv_id[nd] * dim);
        float acc = 0.0f, na = 0.0f, nb = 0.0f;
        for (int j0 = gl; j0 < n4; j0 += B * G) {
            float4 xs[B], ys[B];
#pragma unroll
            for (int u = 0; u < B; u++) {
                // unconditional loads (clamped index), zeros selected after
                const int j = j0 + u * G, jc = min(j, n4 - 1);
                const float4 xa = a[jc], yb = b[jc];
                xs[u] = j < n4 ? xa : z;
                ys[u] = j < n4 ? yb : z;
            }
#pragma unroll
            for (int u = 0; u < B; u++) {
                const float xv[4] = {xs[u].x, xs[u].y, xs[u].z, xs[u].w};
                const float yv[4] = {ys[u].x, ys[u].y, ys[u].z, ys[u].w};
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    if constexpr (M == GTS_METRIC_ANGULAR) {
                        acc = fmaf(xv[e], yv[e], acc);
                        na = fmaf(xv[e], xv[e], na);
                        nb = fmaf(yv[e], yv[e], nb);
                    } else {
                        acc = Acc<M>::step(acc, xv[e], yv[e]);
                    }
                }
            }
        }
        float key;
        if constexpr (M == GTS_METRIC_ANGULAR) {
#pragma unroll
            for (int o = G / 2; o > 0; o >>= 1) {
                acc += __shfl_xor(acc, o, G);
                na += __shfl_xor(na, o, G);
                nb += __shfl_xor(nb, o, G);
            }
            key = angle_deg(acc, na, nb);
        } else {
#pragma unroll
            for (int o = G / 2; o > 0; o >>= 1) acc = Acc<M>::comb(acc, __shfl_xor(acc, o, G));
            key = Acc<M>::fin(acc);
        }
        if (gl == 0) dkey[pos] = __float_as_uint(key);
    }
}

// The same for any dim: LPO lanes per position, 4-byte loads.
template <int M, int LPO>
__global__ void __launch_bounds__(256) k_keys_pos(const float *__restrict__ X0, int dim, const int32_t *perm,
                                                  const int32_t *pos_node, const uint8_t *nflag,
                                                  const int32_t *npiv_id, int64_t s0, int64_t s1, int64_t n,
                                                  uint32_t *dkey)
{
    const int gl = threadIdx.x % LPO;
    const int64_t ng = ((int64_t)gridDim.x * blockDim.x) / LPO;
    for (int64_t pos = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / LPO; pos < n; pos += ng) {
        const int32_t nd = pos_node[pos];
        if (!(nd >= s0 && nd < s1 && is_split(nflag[nd]))) continue;
        const float key = group_dist<M, LPO>(X0 + (int64_t)perm[pos] * dim, X0 + (int64_t)npiv_id[nd] * dim, dim, gl);
        if (gl == 0) dkey[pos] = __float_as_uint(key);
    }
}

// Sorts the range of every splitting node of the level (one workgroup per
// node of nodes s0 .. s1) by (distance key, object ID): the words (key << 32
// | ID) of the range, padded to P (a power of two, 2 x the workgroup size at
// least) with all-ones words, go through a bitonic network in LDS; the
// sorted IDs and keys are written back in place.
__global__ void k_seg_sort(int32_t *perm, uint32_t *dkey, const int32_t *nlid, const int32_t *nsize,
                           const uint8_t *nflag, int64_t s0, int P)
{
    extern __shared__ uint64_t sk[];
    const int64_t nd = s0 + blockIdx.x;
    if (!is_split(nflag[nd])) return;   // the whole workgroup
    const int32_t lid = nlid[nd], size = nsize[nd];
    const int t = threadIdx.x, nt = blockDim.x;
    for (int i = t; i < P; i += nt)
        sk[i] = i < size ? ((uint64_t)dkey[lid + i] << 32) | (uint32_t)perm[lid + i] : ~0ull;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int c = t; c < P / 2; c += nt) {
                // the c-th compare-exchange pair (i, i + j): i has bit j clear
                const int i = ((c & ~(j - 1)) << 1) | (c & (j - 1));
                const uint64_t a = sk[i], b = sk[i + j];
                const bool up = (i & k) == 0;
                if ((a > b) == up) {
                    sk[i] = b;
                    sk[i + j] = a;
                }
            }
            __syncthreads();
        }
    }
    for (int i = t; i < size; i += nt) {
        const uint64_t v = sk[i];
        perm[lid + i] = (int32_t)(uint32_t)v;
        dkey[lid + i] = (uint32_t)(v >> 32);
    }
}

// Keys of the global sort in object-ID order: (segment << 32 | distance key)
// with the segment the node's rank at the level (by_rank) or the start of the
// object's range; an object outside the level's splitting nodes (in a
// finished leaf) is keyed by its offset in its range instead, so the leaf
// keeps its order (the reference's stable sort does, on the node code alone,
// tree.cuh:159-170, 372).
__global__ void k_global_keys(const int32_t *perm, const int32_t *pos_node, const int32_t *nlid,
                              const uint8_t *nflag, const uint32_t *dkey, int64_t s0, int64_t s1, int by_rank,
                              int64_t n, uint64_t *Kid)
{
    for (int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; pos < n;
         pos += (int64_t)gridDim.x * blockDim.x) {
        const int32_t nd = pos_node[pos];
        const bool sp = nd >= s0 && nd < s1 && is_split(nflag[nd]);
        const uint64_t seg = by_rank ? (uint64_t)(nd - s0) : (uint64_t)(uint32_t)nlid[nd];
        Kid[perm[pos]] = (seg << 32) | (sp ? dkey[pos] : (uint32_t)(pos - nlid[nd]));
    }
}

// After the global sort: the sorted IDs become the permutation, the sorted
// keys' distance bits the position-order keys.
__global__ void k_take_sorted(const uint64_t *Ks, const int32_t *Vs, int64_t n, int32_t *perm, uint32_t *dkey)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        perm[i] = Vs[i];
        dkey[i] = (uint32_t)Ks[i];
    }
}

// Edit-distance keys in position order, one thread per position.
__global__ void __launch_bounds__(256) k_keys_str(const uint8_t *__restrict__ S0, int stride,
                                                  const int32_t *len, const int32_t *perm,
                                                  const int32_t *pos_node, const uint8_t *nflag,
                                                  const int32_t *npiv_id, int64_t s0, int64_t s1, int64_t n,
                                                  uint32_t *dkey)
{
    for (int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; pos < n;
         pos += (int64_t)gridDim.x * blockDim.x) {
        const int32_t nd = pos_node[pos];
        if (!(nd >= s0 && nd < s1 && is_split(nflag[nd]))) continue;
        const int32_t x = perm[pos], piv = npiv_id[nd];
        const float key = (float)gts_edit_distance(S0 + (int64_t)x * stride, len[x],
                                                   S0 + (int64_t)piv * stride, len[piv]);
        dkey[pos] = __float_as_uint(key);
    }
}

__global__ void k_split(int32_t *nlid, int32_t *nsize, float *nmin, uint8_t *nflag,
                        const uint32_t *dkey, int64_t s0, int64_t s1, int T, int max_size,
                        int last, int32_t *nsplit_next)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t nd = s0 + t / T;
    int i = (int)(t % T);
    if (nd >= s1 || !is_split(nflag[nd])) return;
    int32_t lid = nlid[nd], size = nsize[nd];
    int32_t avg = size / T;
    int32_t clid = lid + avg * i;
    int32_t csz = (i < T - 1) ? avg : size - (T - 1) * avg;
    int64_t ch = nd * T + 1 + i;
    nlid[ch] = clid;
    nsize[ch] = csz;
    nmin[ch] = __uint_as_float(dkey[clid]);
    bool leaf = csz <= max_size || csz < T || last;
    nflag[ch] = (uint8_t)(1 | (leaf ? 2 : 0));
    if (!leaf) atomicAdd(nsplit_next, 1);
}

__global__ void k_descend(int32_t *pos_node, const int32_t *nlid, const int32_t *nsize,
                          const uint8_t *nflag, int64_t s0, int64_t s1, int T, int64_t n)
{
    for (int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; pos < n;
         pos += (int64_t)gridDim.x * blockDim.x) {
        int32_t nd = pos_node[pos];
        if (nd < s0 || nd >= s1 || !is_split(nflag[nd])) continue;
        int32_t lid = nlid[nd], size = nsize[nd];
        int32_t avg = size / T;
        int c = avg > 0 ? min((int32_t)((pos - lid) / avg), T - 1) : T - 1;
        pos_node[pos] = (int32_t)((int64_t)nd * T + 1 + c);
    }
}

__global__ void k_finalize_vec(const float *__restrict__ X0, int dim, int dp, const int32_t *perm,
                               int64_t n, float *X, float *xn2, int32_t *inv)
{
    // one wave per position
    const int lane = threadIdx.x & 63;
    int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t pos = w; pos < n; pos += nw) {
        int32_t x = perm[pos];
        float s = 0.f;
        for (int j = lane; j < dp; j += 64) {
            float v = j < dim ? X0[(int64_t)x * dim + j] : 0.0f;
            X[pos * dp + j] = v;
            s = fmaf(v, v, s);
        }
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if (lane == 0) {
            inv[x] = (int32_t)pos;
            xn2[pos] = s;
        }
    }
}

// The same for dim % 4 == 0 (dp == dim): 16 lanes per position, four
// positions per wave, 16-byte loads and stores.
__global__ void __launch_bounds__(256) k_finalize_vec4(const float *__restrict__ X0, int dim, const int32_t *perm,
                                                       int64_t n, float *X, float *xn2, int32_t *inv)
{
    constexpr int G = 16;
    const int gl = threadIdx.x % G;
    const int n4 = dim >> 2;
    const int64_t ng = ((int64_t)gridDim.x * blockDim.x) / G;
    for (int64_t pos = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G; pos < n; pos += ng) {
        const int32_t x = perm[pos];
        const float4 *src = (const float4 *)(X0 + (int64_t)x * dim);
        float4 *dst = (float4 *)(X + pos * dim);
        float s = 0.0f;
        for (int j = gl; j < n4; j += G) {
            const float4 v = src[j];
            dst[j] = v;
            s = fmaf(v.x, v.x, s);
            s = fmaf(v.y, v.y, s);
            s = fmaf(v.z, v.z, s);
            s = fmaf(v.w, v.w, s);
        }
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
        if (gl == 0) {
            inv[x] = (int32_t)pos;
            xn2[pos] = s;
        }
    }
}

__global__ void k_finalize_str(const uint8_t *S0, int stride, const int32_t *len0,
                               const int32_t *perm, int64_t n, uint8_t *S, int sstride,
                               int32_t *slen, int32_t *inv)
{
    for (int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; pos < n;
         pos += (int64_t)gridDim.x * blockDim.x) {
        int32_t x = perm[pos];
        int L = len0[x];
        for (int j = 0; j < sstride; j++) S[pos * sstride + j] = j < L ? S0[(int64_t)x * stride + j] : 0;
        slen[pos] = L;
        inv[x] = (int32_t)pos;
    }
}

// packed child tables of every node (zero masks for leaves and absent nodes)
__global__ void k_child_tables(const int32_t *nsize, const float *nmin, const uint8_t *nflag,
                               int64_t nodes, int T, float *ring, uint32_t *cne, uint32_t *cleaf)
{
    for (int64_t P = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; P < nodes;
         P += (int64_t)gridDim.x * blockDim.x) {
        uint32_t ne = 0, lf = 0;
        float *r = ring + P * (T + 1);
        const bool split = is_split(nflag[P]);
        for (int c = 0; c < T; c++) {
            const int64_t ch = P * T + 1 + c;
            const bool ok = split && ch < nodes;
            r[c] = ok ? nmin[ch] : 0.0f;
            if (ok && nsize[ch] > 0) ne |= 1u << c;
            if (ok && (nflag[ch] & 2)) lf |= 1u << c;
        }
        r[T] = INFINITY;
        cne[P] = ne;
        cleaf[P] = lf;
    }
}

__global__ void k_piv_pos(const int32_t *npiv_id, const int32_t *inv, int32_t *npiv, int64_t nn)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn;
         i += (int64_t)gridDim.x * blockDim.x)
        npiv[i] = npiv_id[i] >= 0 ? inv[npiv_id[i]] : -1;
}

template <class T>
T *dmalloc(size_t count)
{
    T *p = nullptr;
    if (count == 0) count = 1;
    if (hipMalloc((void **)&p, count * sizeof(T)) != hipSuccess)
        fail(GTS_E_OOM, "device allocation of " + std::to_string(count * sizeof(T)) + " bytes");
    return p;
}

template <class T>
void dfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

}  // namespace

void release_tree(gts_index *idx)
{
    dfree(idx->X); dfree(idx->xn2); dfree(idx->Xhi); dfree(idx->Xlo); dfree(idx->S);
    idx->dk = 0;
    idx->xn2_max_ok = false;
    idx->par_level = -1;
    dfree(idx->slen); dfree(idx->perm); dfree(idx->inv); dfree(idx->nlid); dfree(idx->nsize);
    dfree(idx->npiv); dfree(idx->nmin); dfree(idx->nflag); dfree(idx->ring); dfree(idx->cne);
    dfree(idx->cleaf);
    // the radius network's feature pivots are tree positions of this tree, and
    // the plan and schedule choices were made for it
    idx->model.trained = false;
    idx->rp_n = -1;
    idx->plan_memo.clear();
    idx->dense_sched.clear();
    idx->lowdim_sched.clear();
    graph_reset(idx);
    cells_release(idx);
}

void build_tree(gts_index *idx, const gts_dataset *src)
{
    hipStream_t st = idx->st;
    const gts_dataset &D = src ? *src : idx->data;
    const int64_t n = D.n;
    const int T = idx->T;
    release_tree(idx);
    idx->n = n;
    idx->metric = D.metric;
    const bool is_str = D.metric == GTS_METRIC_EDIT;
    idx->dim = is_str ? 0 : D.dim;
    idx->dp = is_str ? 0 : ((D.dim + 3) / 4) * 4;
    idx->em = err_model(D.metric, D.dim);

    const int h = tree_height(n, T, idx->max_size, idx->max_h);
    idx->height = h;
    idx->level_start.assign(h + 1, 0);
    int64_t nodes = 0, pw = 1;
    for (int l = 0; l < h; l++) {
        idx->level_start[l] = nodes;
        nodes += pw;
        if (nodes > (int64_t)INT32_MAX) fail(GTS_E_ARG, "tree too large for 32-bit node IDs");
        pw *= T;
    }
    idx->level_start[h] = nodes;
    idx->num_nodes = nodes;
    idx->pnum.assign(h + 1, 0);

    idx->nlid = dmalloc<int32_t>(nodes);
    idx->nsize = dmalloc<int32_t>(nodes);
    idx->npiv = dmalloc<int32_t>(nodes);
    idx->nmin = dmalloc<float>(nodes);
    idx->nflag = dmalloc<uint8_t>(nodes);
    idx->perm = dmalloc<int32_t>(n);
    idx->inv = dmalloc<int32_t>(n);
    GTS_HIP(hipMemsetAsync(idx->nlid, 0, nodes * sizeof(int32_t), st));
    GTS_HIP(hipMemsetAsync(idx->nsize, 0, nodes * sizeof(int32_t), st));
    GTS_HIP(hipMemsetAsync(idx->nmin, 0, nodes * sizeof(float), st));
    GTS_HIP(hipMemsetAsync(idx->nflag, 0, nodes, st));
    int32_t *npiv_id = dmalloc<int32_t>(nodes);
    fill_i32(npiv_id, -1, nodes, st);

    // root {size = n, lid = 0} (initIndexData, tree.cuh:299-318)
    const bool root_leaf = n <= idx->max_size || n < T || h == 1;
    uint8_t rf = (uint8_t)(1 | (root_leaf ? 2 : 0));
    int32_t rs = (int32_t)n;
    GTS_HIP(hipMemcpyAsync(idx->nflag, &rf, 1, hipMemcpyHostToDevice, st));
    GTS_HIP(hipMemcpyAsync(idx->nsize, &rs, sizeof(int32_t), hipMemcpyHostToDevice, st));

    // original-order objects on the device
    float *X0 = nullptr;
    uint8_t *S0 = nullptr;
    int32_t *L0 = nullptr;
    if (is_str) {
        S0 = dmalloc<uint8_t>((size_t)n * D.stride);
        L0 = dmalloc<int32_t>(n);
        GTS_HIP(hipMemcpyAsync(S0, D.str.data(), (size_t)n * D.stride, hipMemcpyHostToDevice, st));
        GTS_HIP(hipMemcpyAsync(L0, D.len.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
    } else {
        X0 = dmalloc<float>((size_t)n * D.dim);
        GTS_HIP(hipMemcpyAsync(X0, D.vec.data(), (size_t)n * D.dim * sizeof(float),
                               hipMemcpyHostToDevice, st));
    }

    fill_iota_i32(idx->perm, n, st);
    int32_t *pos_node = dmalloc<int32_t>(n);
    GTS_HIP(hipMemsetAsync(pos_node, 0, n * sizeof(int32_t), st));
    uint32_t *dkey = dmalloc<uint32_t>(n);
    uint64_t *K0 = nullptr, *K1 = nullptr;
    int32_t *V0 = nullptr, *V1 = nullptr;
    int32_t *dcount = dmalloc<int32_t>(1);
    int hb = 1;
    while (hb < 63 && ((int64_t)1 << hb) < n) hb++;
    std::vector<int64_t> maxsplit;
    std::vector<char> allsplit;
    level_plan(n, T, idx->max_size, h, maxsplit, allsplit);

    int64_t splitting = root_leaf ? 0 : 1;
    for (int l = 0; l + 1 < h && splitting > 0; l++) {
        const int64_t s0 = idx->level_start[l], s1 = idx->level_start[l + 1];
        idx->pnum[l + 1] = splitting;
        k_pivots<<<grid_for(s1 - s0, 256), 256, 0, st>>>(idx->perm, idx->nlid, idx->nsize,
                                                         idx->nflag, npiv_id, s0, s1);
        GTS_HIP(hipGetLastError());
        // distance keys of the level's splitting nodes, in position order
        if (is_str) {
            k_keys_str<<<grid_for(n, 256, 65536), 256, 0, st>>>(S0, D.stride, L0, idx->perm, pos_node, idx->nflag,
                                                                npiv_id, s0, s1, n, dkey);
        } else {
            const int dim = D.dim;
            GTS_METRIC_SWITCH(D.metric, {
                if (dim % 4 == 0)
                    k_keys_pos4<M_><<<grid_for(n * 16, 256, 65536), 256, 0, st>>>(
                        X0, dim, idx->perm, pos_node, idx->nflag, npiv_id, s0, s1, n, dkey);
                else if (dim <= 8)
                    k_keys_pos<M_, 1><<<grid_for(n, 256, 65536), 256, 0, st>>>(
                        X0, dim, idx->perm, pos_node, idx->nflag, npiv_id, s0, s1, n, dkey);
                else if (dim <= 128)
                    k_keys_pos<M_, 16><<<grid_for(n * 16, 256, 65536), 256, 0, st>>>(
                        X0, dim, idx->perm, pos_node, idx->nflag, npiv_id, s0, s1, n, dkey);
                else
                    k_keys_pos<M_, 64><<<grid_for(n * 64, 256, 65536), 256, 0, st>>>(
                        X0, dim, idx->perm, pos_node, idx->nflag, npiv_id, s0, s1, n, dkey);
            });
        }
        GTS_HIP(hipGetLastError());
        // each splitting node's range sorted by (key, object ID)
        if (maxsplit[l] <= LDS_SORT_MAX) {
            int P = 64;
            while (P < maxsplit[l]) P <<= 1;
            const int nt = std::min(1024, P / 2);
            const size_t lds = (size_t)P * 8;
            if (lds > 64 * 1024)
                GTS_HIP(hipFuncSetAttribute((const void *)k_seg_sort, hipFuncAttributeMaxDynamicSharedMemorySize,
                                            (int)lds));
            k_seg_sort<<<(unsigned)(s1 - s0), nt, lds, st>>>(idx->perm, dkey, idx->nlid, idx->nsize, idx->nflag,
                                                            s0, P);
        } else {
            if (!K0) {
                K0 = dmalloc<uint64_t>(n);
                K1 = dmalloc<uint64_t>(n);
                V0 = dmalloc<int32_t>(n);
                V1 = dmalloc<int32_t>(n);
            }
            // the segment: the node's rank at the level when every object
            // lies in one of its splitting nodes, else the range start
            const bool by_rank = allsplit[l] != 0;
            int sb = 0;
            while (by_rank && ((int64_t)1 << sb) < s1 - s0) sb++;
            const int bits = 32 + (by_rank ? sb : hb);
            k_global_keys<<<grid_for(n, 256, 65536), 256, 0, st>>>(idx->perm, pos_node, idx->nlid, idx->nflag, dkey,
                                                                   s0, s1, by_rank ? 1 : 0, n, K0);
            GTS_HIP(hipGetLastError());
            fill_iota_i32(V0, n, st);
            uint64_t *Ks;
            int32_t *Vs;
            radix_sort_u64(K0, V0, K1, V1, n, bits, nullptr, 0, st, &Ks, &Vs);
            k_take_sorted<<<grid_for(n, 256, 65536), 256, 0, st>>>(Ks, Vs, n, idx->perm, dkey);
        }
        GTS_HIP(hipGetLastError());
        GTS_HIP(hipMemsetAsync(dcount, 0, sizeof(int32_t), st));
        const int last = (l + 1 == h - 1) ? 1 : 0;
        k_split<<<grid_for((s1 - s0) * T, 256), 256, 0, st>>>(idx->nlid, idx->nsize, idx->nmin,
                                                              idx->nflag, dkey, s0, s1, T,
                                                              idx->max_size, last, dcount);
        GTS_HIP(hipGetLastError());
        k_descend<<<grid_for(n, 256, 65536), 256, 0, st>>>(pos_node, idx->nlid, idx->nsize,
                                                           idx->nflag, s0, s1, T, n);
        GTS_HIP(hipGetLastError());
        int32_t next = 0;
        GTS_HIP(hipMemcpyAsync(&next, dcount, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        GTS_HIP(hipStreamSynchronize(st));
        splitting = next;
    }

    // tree-order copies of the objects, inverse permutation, pivot positions
    if (is_str) {
        int mx = 1;
        for (int32_t L : D.len) mx = L > mx ? L : mx;
        idx->sstride = ((mx + 15) / 16) * 16;
        idx->S = dmalloc<uint8_t>((size_t)n * idx->sstride);
        idx->slen = dmalloc<int32_t>(n);
        k_finalize_str<<<grid_for(n, 256, 65536), 256, 0, st>>>(S0, D.stride, L0, idx->perm, n,
                                                                idx->S, idx->sstride, idx->slen,
                                                                idx->inv);
    } else {
        idx->X = dmalloc<float>((size_t)n * idx->dp);
        idx->xn2 = dmalloc<float>(n);
        if (D.dim % 4 == 0)
            k_finalize_vec4<<<grid_for(n * 16, 256, 65536), 256, 0, st>>>(X0, D.dim, idx->perm, n, idx->X,
                                                                          idx->xn2, idx->inv);
        else
            k_finalize_vec<<<grid_for(n * 64, 256, 65536), 256, 0, st>>>(X0, D.dim, idx->dp, idx->perm,
                                                                         n, idx->X, idx->xn2, idx->inv);
    }
    GTS_HIP(hipGetLastError());
    k_piv_pos<<<grid_for(nodes, 256, 65536), 256, 0, st>>>(npiv_id, idx->inv, idx->npiv, nodes);
    GTS_HIP(hipGetLastError());
    idx->ring = dmalloc<float>((size_t)nodes * (T + 1));
    idx->cne = dmalloc<uint32_t>(nodes);
    idx->cleaf = dmalloc<uint32_t>(nodes);
    k_child_tables<<<grid_for(nodes, 256, 65536), 256, 0, st>>>(idx->nsize, idx->nmin, idx->nflag, nodes, T,
                                                                idx->ring, idx->cne, idx->cleaf);
    GTS_HIP(hipGetLastError());
    GTS_HIP(hipStreamSynchronize(st));
    dfree(npiv_id); dfree(pos_node); dfree(dkey); dfree(K0); dfree(K1); dfree(V0); dfree(V1); dfree(dcount);
    dfree(X0); dfree(S0); dfree(L0);
}

}  // namespace gts
