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
// double key d/INFI_DIS + node*DIS_CODE (tree.cuh:155, 372).  Here the key is
// exact: (start position of the object's current range) << 32 | FP32 distance
// bits, sorted by a stable LSD radix sort whose input is laid out in object-ID
// order, so ties inside a node fall in ID order.  Objects of finished leaves
// keep their range (distance key 0).
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

namespace {

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

template <int M, int LPO>
__global__ void __launch_bounds__(256) k_keys_vec(const float *__restrict__ X0, int dim,
                                                  const int32_t *perm, const int32_t *pos_node,
                                                  const int32_t *nlid, const uint8_t *nflag,
                                                  const int32_t *npiv_id, int64_t s0, int64_t s1,
                                                  int64_t n, uint64_t *Kid)
{
    const int gl = threadIdx.x % LPO;
    int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / LPO;
    int64_t ng = ((int64_t)gridDim.x * blockDim.x) / LPO;
    for (int64_t pos = g; pos < n; pos += ng) {
        int32_t nd = pos_node[pos];
        int32_t x = perm[pos];
        float key = 0.0f;
        if (nd >= s0 && nd < s1 && is_split(nflag[nd])) {
            int32_t piv = npiv_id[nd];
            key = group_dist<M, LPO>(X0 + (int64_t)x * dim, X0 + (int64_t)piv * dim, dim, gl);
        }
        if (gl == 0) Kid[x] = ((uint64_t)(uint32_t)nlid[nd] << 32) | __float_as_uint(key);
    }
}

__global__ void __launch_bounds__(256) k_keys_str(const uint8_t *__restrict__ S0, int stride,
                                                  const int32_t *len, const int32_t *perm,
                                                  const int32_t *pos_node, const int32_t *nlid,
                                                  const uint8_t *nflag, const int32_t *npiv_id,
                                                  int64_t s0, int64_t s1, int64_t n, uint64_t *Kid)
{
    for (int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; pos < n;
         pos += (int64_t)gridDim.x * blockDim.x) {
        int32_t nd = pos_node[pos];
        int32_t x = perm[pos];
        float key = 0.0f;
        if (nd >= s0 && nd < s1 && is_split(nflag[nd])) {
            int32_t piv = npiv_id[nd];
            key = (float)gts_edit_distance(S0 + (int64_t)x * stride, len[x],
                                           S0 + (int64_t)piv * stride, len[piv]);
        }
        Kid[x] = ((uint64_t)(uint32_t)nlid[nd] << 32) | __float_as_uint(key);
    }
}

__global__ void k_split(int32_t *nlid, int32_t *nsize, float *nmin, uint8_t *nflag,
                        const uint64_t *Ks, int64_t s0, int64_t s1, int T, int max_size,
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
    nmin[ch] = __uint_as_float((uint32_t)Ks[clid]);
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

__global__ void k_take_vals(const uint64_t *Ks, uint32_t *keybits, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        keybits[i] = (uint32_t)Ks[i];
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

void build_tree(gts_index *idx)
{
    hipStream_t st = idx->st;
    const gts_dataset &D = idx->data;
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
    uint64_t *K0 = dmalloc<uint64_t>(n), *K1 = dmalloc<uint64_t>(n);
    int32_t *V0 = dmalloc<int32_t>(n), *V1 = dmalloc<int32_t>(n);
    int32_t *dcount = dmalloc<int32_t>(1);
    int hb = 1;
    while (hb < 63 && ((int64_t)1 << hb) < n) hb++;
    const int bits = 32 + hb;

    int64_t splitting = root_leaf ? 0 : 1;
    for (int l = 0; l + 1 < h && splitting > 0; l++) {
        const int64_t s0 = idx->level_start[l], s1 = idx->level_start[l + 1];
        idx->pnum[l + 1] = splitting;
        k_pivots<<<grid_for(s1 - s0, 256), 256, 0, st>>>(idx->perm, idx->nlid, idx->nsize,
                                                         idx->nflag, npiv_id, s0, s1);
        GTS_HIP(hipGetLastError());
        if (is_str) {
            k_keys_str<<<grid_for(n, 256, 65536), 256, 0, st>>>(S0, D.stride, L0, idx->perm,
                                                                pos_node, idx->nlid, idx->nflag,
                                                                npiv_id, s0, s1, n, K0);
        } else {
            const int dim = D.dim;
            GTS_METRIC_SWITCH(D.metric, {
                if (dim <= 8)
                    k_keys_vec<M_, 1><<<grid_for(n, 256, 65536), 256, 0, st>>>(
                        X0, dim, idx->perm, pos_node, idx->nlid, idx->nflag, npiv_id, s0, s1, n, K0);
                else if (dim <= 128)
                    k_keys_vec<M_, 16><<<grid_for(n * 16, 256, 65536), 256, 0, st>>>(
                        X0, dim, idx->perm, pos_node, idx->nlid, idx->nflag, npiv_id, s0, s1, n, K0);
                else
                    k_keys_vec<M_, 64><<<grid_for(n * 64, 256, 65536), 256, 0, st>>>(
                        X0, dim, idx->perm, pos_node, idx->nlid, idx->nflag, npiv_id, s0, s1, n, K0);
            });
        }
        GTS_HIP(hipGetLastError());
        fill_iota_i32(V0, n, st);
        uint64_t *Ks;
        int32_t *Vs;
        radix_sort_u64(K0, V0, K1, V1, n, bits, nullptr, 0, st, &Ks, &Vs);
        GTS_HIP(hipMemcpyAsync(idx->perm, Vs, n * sizeof(int32_t), hipMemcpyDeviceToDevice, st));
        GTS_HIP(hipMemsetAsync(dcount, 0, sizeof(int32_t), st));
        const int last = (l + 1 == h - 1) ? 1 : 0;
        k_split<<<grid_for((s1 - s0) * T, 256), 256, 0, st>>>(idx->nlid, idx->nsize, idx->nmin,
                                                              idx->nflag, Ks, s0, s1, T,
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
    dfree(npiv_id); dfree(pos_node); dfree(K0); dfree(K1); dfree(V0); dfree(V1); dfree(dcount);
    dfree(X0); dfree(S0); dfree(L0);
}

}  // namespace gts
