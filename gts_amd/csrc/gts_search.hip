// Two-stage concurrent range and kNN search on gfx950.
//
// Replaces searchIndexRnnV2 / searchIndexKnnV2 (search_v2.cuh:780-1228).
// Stage 1 walks the tree level by level with a sparse frontier of (query,
// node) pairs instead of the reference's dense T^l x q flag tables
// (search_v2.cuh:40-196): for every surviving pair the distance from the query
// to the node's pivot is computed once (the reference recomputes it for each
// of the T children, search_v2.cuh:63-137), children are kept when their ring
// lower bound max(min_dis_i - dq, dq - min_dis_{i+1}) does not exceed the
// query's bound (search_v2.cuh:139-148), internal children form the next
// frontier and surviving leaf children become (query, parent, leaf-mask)
// records.  For kNN the bound is the k-th smallest pivot distance over the
// query's surviving parents from level 3 on (getDisPQ + sort + updateDisK,
// search_v2.cuh:645-777), found by a radix select instead of a global sort.
// Stage 2 groups the records by parent node and verifies them in tiles of
// 64 queries x 64 objects (dataProcessRnn / dataProcessKnn,
// search_v2.cuh:319-544, which use one 512-thread block per pair with 20
// active lanes).  FP32 distances are exact up to a stated error model; pairs
// inside the band around the radius are re-decided in FP64, and the kNN
// selection re-ranks its boundary candidates in FP64, so result sets equal the
// exact ones.
#include <algorithm>

#include "gts_cells.h"
#include "gts_dist.h"
#include "gts_index.h"
#include "gts_search.h"
#include "gts_mfma.h"
#include "gts_verify2.h"
#include "gts_wave.h"

namespace gts {

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int QT = 64;    // queries per tile
constexpr int RT = 64;    // objects per tile
constexpr int DK = 32;    // dimensions per LDS stage
constexpr int LDP = 68;   // LDS row pitch (floats), keeps 16-byte alignment

unsigned grid_for(int64_t n, int block, int64_t cap = 1 << 20)
{
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

__device__ __forceinline__ bool leaf_flag(uint8_t f) { return (f & 2) != 0; }

// Range [lid, lid + size) of child c of node P; P = -1 is the root itself.
__device__ __forceinline__ void child_range(const DevTree &t, int32_t P, int c, int32_t &lid,
                                            int32_t &sz)
{
    if (P < 0) {
        lid = 0;
        sz = (int32_t)t.n;
    } else {
        int64_t ch = (int64_t)P * t.T + 1 + c;
        lid = t.nlid[ch];
        sz = t.nsize[ch];
    }
}

// ---------------------------------------------------------------- stage 1
template <int M, int LPO>
__global__ void __launch_bounds__(256) k_pivot_dist(DevTree t, const int32_t *fq,
                                                    const int32_t *fn, int64_t nf,
                                                    const int32_t *qpos, float *dq)
{
    const int gl = threadIdx.x % LPO;
    int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / LPO;
    int64_t ng = ((int64_t)gridDim.x * blockDim.x) / LPO;
    for (int64_t i = g; i < nf; i += ng) {
        int32_t qp = qpos[fq[i]];
        int32_t pp = t.npiv[fn[i]];
        float d = group_dist<M, LPO>(t.X + (int64_t)qp * t.dp, t.X + (int64_t)pp * t.dp, t.dim, gl);
        if (gl == 0) dq[i] = d;
    }
}

__global__ void __launch_bounds__(256) k_pivot_dist_edit(DevTree t, const int32_t *fq,
                                                         const int32_t *fn, int64_t nf,
                                                         const int32_t *qpos, float *dq)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nf;
         i += (int64_t)gridDim.x * blockDim.x) {
        int32_t qp = qpos[fq[i]];
        int32_t pp = t.npiv[fn[i]];
        dq[i] = (float)gts_edit_distance(t.S + (int64_t)qp * t.sstride, t.slen[qp],
                                         t.S + (int64_t)pp * t.sstride, t.slen[pp]);
    }
}

// k-th smallest of the query's pivot distances (radix select on FP32 bits);
// tightens the query's bound (updateDisK, search_v2.cuh:755-777).
__global__ void __launch_bounds__(256) k_knn_bound(const float *dq, const int64_t *qoff,
                                                   const int32_t *kq, int64_t pnum_l, int nq,
                                                   float *bnd, ErrModel em)
{
    __shared__ uint32_t hist[256];
    __shared__ uint32_t sh_prefix;
    __shared__ int sh_k;
    const int q = blockIdx.x;
    if (q >= nq) return;
    const int64_t s = qoff[q], e = qoff[q + 1];
    const int k = kq[q];
    if (pnum_l < k || e - s < k) return;
    uint32_t prefix = 0, himask = 0;
    int kk = k;
    for (int shift = 24; shift >= 0; shift -= 8) {
        hist[threadIdx.x] = 0;
        __syncthreads();
        for (int64_t i = s + threadIdx.x; i < e; i += blockDim.x) {
            uint32_t b = __float_as_uint(dq[i]);
            if ((b & himask) == prefix) atomicAdd(&hist[(b >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t cum = 0;
            int dgt = 0;
            for (; dgt < 256; dgt++) {
                if (cum + hist[dgt] >= (uint32_t)kk) break;
                cum += hist[dgt];
            }
            sh_k = kk - (int)cum;
            sh_prefix = prefix | ((uint32_t)dgt << shift);
        }
        __syncthreads();
        kk = sh_k;
        prefix = sh_prefix;
        himask |= 255u << shift;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        float kth = __uint_as_float(prefix);
        float up = kth + em.rel * kth * 2.0f + em.abs;
        if (up < bnd[q]) bnd[q] = up;
    }
}

// Ring tests (search_v2.cuh:139-148) of all T children of one frontier pair
// from the parent's packed child table (ring keys, non-empty and leaf masks),
// with a margin covering the FP32 error of dq and of the ring keys; returns the
// mask of kept children.
__device__ __forceinline__ uint32_t kept_children(const DevTree &t, int32_t P, float d, float b,
                                                  ErrModel em)
{
    const float *ring = t.ring + (int64_t)P * (t.T + 1);
    const uint32_t ne = t.cne[P];
    const float bm = isinf(b) ? 0.0f : b;
    uint32_t km = 0;
    float lo = ring[0];
    for (int c = 0; c < t.T; c++) {
        const float hi = ring[c + 1];
        float lb = lo - d;
        float hv = 0.0f;
        if (c < t.T - 1) {
            hv = hi;
            lb = fmaxf(lb, d - hi);
        }
        const float margin = 2.0f * em.rel * (d + lo + hv + bm) + 2.0f * em.abs;
        if (((ne >> c) & 1u) && !(lb > b + margin)) km |= 1u << c;
        lo = hi;
    }
    return km;
}

// Ring tests of the frontier pairs (nodeProcessRnn / nodeProcessKnn,
// search_v2.cuh:40-196): per pair the mask of kept children, the number of
// kept inner children and whether a kept child is a leaf.
__global__ void __launch_bounds__(256) k_children_count(DevTree t, const int32_t *fq,
                                                        const int32_t *fn, const float *dq,
                                                        int64_t nf, const float *bnd, ErrModel em,
                                                        int32_t *cnt_int, int32_t *cnt_leaf,
                                                        uint32_t *kmask)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nf;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t q = fq[i], P = fn[i];
        const uint32_t km = kept_children(t, P, dq[i], bnd[q], em);
        const uint32_t lf = t.cleaf[P];
        const int ni = __popc(km & ~lf);
        cnt_int[i] = ni;
        cnt_leaf[i] = (km & lf) != 0;
        kmask[i] = km;
    }
}

// The next frontier's query offsets.  The frontier is in query order (the
// root pairs are, and each pair's children are written at the exclusive scan
// of the kept counts), so query q's next pairs start where the scan stands at
// its first pair now: qoff[q] <- off_int[qoff[q]] (nn past the last pair).
__global__ void k_next_qoff(int64_t *qoff, int64_t nq, const int64_t *off_int, int64_t nf, int64_t nn)
{
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q <= nq; q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t o = qoff[q];
        qoff[q] = o < nf ? off_int[o] : nn;
    }
}

__global__ void __launch_bounds__(256) k_children_write(DevTree t, const int32_t *fq,
                                                        const int32_t *fn, int64_t nf,
                                                        const int64_t *off_int,
                                                        const int64_t *off_leaf,
                                                        const uint32_t *kmask, int32_t *nq_out,
                                                        int32_t *nn_out, int64_t rec_base,
                                                        int32_t *rec_q, int32_t *rec_n,
                                                        uint32_t *rec_m)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nf;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t q = fq[i], P = fn[i];
        const uint32_t km = kmask[i], lf = t.cleaf[P];
        uint32_t inner = km & ~lf;
        int64_t o = off_int[i];
        const int64_t c0 = (int64_t)P * t.T + 1;
        while (inner) {
            const int c = __ffs(inner) - 1;
            inner &= inner - 1;
            nq_out[o] = q;
            nn_out[o] = (int32_t)(c0 + c);
            o++;
        }
        const uint32_t m = km & lf;
        if (m) {
            const int64_t r = rec_base + off_leaf[i];
            rec_q[r] = q;
            rec_n[r] = P;
            rec_m[r] = m;
        }
    }
}

__global__ void k_init_frontier(int32_t *fq, int32_t *fn, int32_t *rq, int32_t *rn, uint32_t *rm,
                                int nq, int root_leaf, int64_t *qoff)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x) {
        if (root_leaf) {
            rq[i] = i;
            rn[i] = -1;
            rm[i] = 1u;
        } else {
            fq[i] = i;
            fn[i] = 0;
        }
        qoff[i] = i;
        if (i == nq - 1) qoff[nq] = nq;
    }
}

// ------------------------------------------------------- group by node
__global__ void k_node_hist(const int32_t *rn, int64_t nr, int32_t *cnt)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nr;
         i += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&cnt[rn[i] + 1], 1);
}

__global__ void k_node_scatter(const int32_t *rq, const int32_t *rn, const uint32_t *rm, int64_t nr,
                               unsigned long long *cursor, int32_t *gq, int32_t *gn, uint32_t *gm)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nr;
         i += (int64_t)gridDim.x * blockDim.x) {
        int32_t s = rn[i] + 1;
        unsigned long long o = atomicAdd(&cursor[s], 1ull);
        gq[o] = rq[i];
        gn[o] = rn[i];
        gm[o] = rm[i];
    }
}

__global__ void k_tiles_per_node(const int32_t *cnt, int64_t ns, int32_t *tiles)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ns;
         i += (int64_t)gridDim.x * blockDim.x)
        tiles[i] = (cnt[i] + QT - 1) / QT;
}

// ---------------------------------------------------------------- stage 2
struct VerifyOut {
    unsigned long long *counts;  // range: per query
    uint64_t *cand;              // knn: per query [cap]
    int32_t *ccount;
    int cap;
    int32_t *band_q, *band_p;    // range band pairs
    unsigned long long *band_n;
    int64_t band_cap;
    uint64_t *res_k;             // range ids: (q << 32 | id)
    float *res_d;
    unsigned long long *res_n;
    int64_t res_cap;
    unsigned long long *verified;
};

enum { MODE_COUNT = 0, MODE_IDS = 1, MODE_KNN = 2 };

__device__ __forceinline__ int64_t upper_bound_i64(const int64_t *a, int64_t n, int64_t v)
{
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (a[mid] <= v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

template <int M, int MODE>
__global__ void __launch_bounds__(256) k_verify(DevTree t, const int32_t *gq, const uint32_t *gm,
                                                const int64_t *node_off, const int64_t *tile_off,
                                                int64_t nslots, const int32_t *qpos,
                                                const float *bnd, ErrModel em, VerifyOut out)
{
    __shared__ __attribute__((aligned(16))) float Qs[DK][LDP];
    __shared__ __attribute__((aligned(16))) float Xs[DK][LDP];
    __shared__ int32_t s_q[QT], s_qp[QT];
    __shared__ uint32_t s_m[QT];
    __shared__ float s_b[QT];
    __shared__ uint32_t s_union;
    __shared__ int32_t s_cnt[QT];
    __shared__ unsigned long long s_base[QT];
    __shared__ uint32_t s_tile_n;
    __shared__ unsigned long long s_tile_base;

    const int tid = threadIdx.x;
    const int64_t tile = blockIdx.x;
    const int64_t slot = upper_bound_i64(tile_off, nslots + 1, tile) - 1;
    const int64_t chunk = tile - tile_off[slot];
    const int64_t rstart = node_off[slot] + chunk * QT;
    const int rcount = (int)min((int64_t)QT, node_off[slot + 1] - rstart);
    const int32_t P = (int32_t)slot - 1;

    if (tid == 0) s_union = 0;
    __syncthreads();
    if (tid < QT) {
        if (tid < rcount) {
            int32_t q = gq[rstart + tid];
            s_q[tid] = q;
            s_qp[tid] = qpos[q];
            s_m[tid] = gm[rstart + tid];
            s_b[tid] = bnd[q];
            atomicOr(&s_union, gm[rstart + tid]);
        } else {
            s_q[tid] = -1;
            s_qp[tid] = 0;
            s_m[tid] = 0;
            s_b[tid] = -1.0f;
        }
    }
    __syncthreads();
    const uint32_t umask = s_union;
    const int ty = tid >> 4, tx = tid & 15;
    const int lrow = tid >> 2, lk = (tid & 3) * 8;
    const int dp = t.dp;
    unsigned long long nver = 0;
    const int nchild = P < 0 ? 1 : t.T;
    const int32_t plid = P < 0 ? 0 : t.nlid[P];
    const int32_t pavg = P < 0 ? 0 : t.nsize[P] / t.T;

    // Rows of consecutive leaf children in the tile's mask are contiguous
    // positions, so each run is tiled as one range (leaves hold ~10 rows).
    for (int c = 0; c < nchild; c++) {
        if (!((umask >> c) & 1u)) continue;
        int cend = c;
        while (cend + 1 < nchild && ((umask >> (cend + 1)) & 1u)) cend++;
        int32_t clid, csz, elid, esz;
        child_range(t, P, c, clid, csz);
        child_range(t, P, cend, elid, esz);
        const int32_t run = elid + esz - clid;
        c = cend;
        for (int r0 = 0; r0 < run; r0 += RT) {
            float acc[4][4];
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 4; b++) acc[a][b] = 0.0f;
            for (int k0 = 0; k0 < dp; k0 += DK) {
                const int kw = min(DK, dp - k0);
                // stage query rows and object rows (dimension-major in LDS)
                {
                    float4 qa = make_float4(0, 0, 0, 0), qb = qa, xa = qa, xb = qa;
                    const bool qv = lrow < rcount;
                    const bool xv = r0 + lrow < run;
                    const float *qr = t.X + (int64_t)s_qp[lrow] * dp + k0 + lk;
                    const float *xr = t.X + (int64_t)(clid + r0 + lrow) * dp + k0 + lk;
                    if (lk < kw) {
                        if (qv) qa = *(const float4 *)qr;
                        if (xv) xa = *(const float4 *)xr;
                    }
                    if (lk + 4 < kw) {
                        if (qv) qb = *(const float4 *)(qr + 4);
                        if (xv) xb = *(const float4 *)(xr + 4);
                    }
                    Qs[lk + 0][lrow] = qa.x; Qs[lk + 1][lrow] = qa.y;
                    Qs[lk + 2][lrow] = qa.z; Qs[lk + 3][lrow] = qa.w;
                    Qs[lk + 4][lrow] = qb.x; Qs[lk + 5][lrow] = qb.y;
                    Qs[lk + 6][lrow] = qb.z; Qs[lk + 7][lrow] = qb.w;
                    Xs[lk + 0][lrow] = xa.x; Xs[lk + 1][lrow] = xa.y;
                    Xs[lk + 2][lrow] = xa.z; Xs[lk + 3][lrow] = xa.w;
                    Xs[lk + 4][lrow] = xb.x; Xs[lk + 5][lrow] = xb.y;
                    Xs[lk + 6][lrow] = xb.z; Xs[lk + 7][lrow] = xb.w;
                }
                __syncthreads();
                for (int kk = 0; kk < kw; kk++) {
                    float4 qv4 = *(const float4 *)&Qs[kk][ty * 4];
                    float4 xv4 = *(const float4 *)&Xs[kk][tx * 4];
                    float qa[4] = {qv4.x, qv4.y, qv4.z, qv4.w};
                    float xa[4] = {xv4.x, xv4.y, xv4.z, xv4.w};
#pragma unroll
                    for (int a = 0; a < 4; a++)
#pragma unroll
                        for (int b = 0; b < 4; b++) acc[a][b] = Acc<M>::step(acc[a][b], qa[a], xa[b]);
                }
                __syncthreads();
            }
            // epilogue: decide each pair
            if (tid < QT) s_cnt[tid] = 0;
            if (tid == 0) s_tile_n = 0;
            __syncthreads();
            int slot_local[4][4];
            uint32_t hitmask = 0, bandmask = 0;
#pragma unroll
            for (int a = 0; a < 4; a++) {
                const int i = ty * 4 + a;
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    const int j = tx * 4 + b;
                    slot_local[a][b] = 0;
                    const int32_t pj = clid + r0 + j;
                    const int cj = P < 0 ? 0 : (pavg > 0 ? min((pj - plid) / pavg, t.T - 1) : t.T - 1);
                    const bool valid = i < rcount && r0 + j < run && ((s_m[i] >> cj) & 1u) &&
                                       !hidden(t, pj, s_q[i]);
                    if (!valid) continue;
                    nver++;
                    const float d = tile_fin<M>(t, acc[a][b], s_qp[i], pj);
                    const float e = em.rel * d + em.abs;
                    const float bq = s_b[i];
                    if (MODE == MODE_KNN) {
                        if (d - e <= bq) {
                            hitmask |= 1u << (a * 4 + b);
                            slot_local[a][b] = atomicAdd(&s_cnt[i], 1);
                        }
                    } else {
                        if (d + e <= bq) {
                            hitmask |= 1u << (a * 4 + b);
                            if (MODE == MODE_IDS) slot_local[a][b] = (int)atomicAdd(&s_tile_n, 1u);
                            atomicAdd(&s_cnt[i], 1);
                        } else if (d - e <= bq) {
                            bandmask |= 1u << (a * 4 + b);
                        }
                    }
                }
            }
            __syncthreads();
            if (MODE == MODE_KNN) {
                if (tid < QT && s_cnt[tid] > 0)
                    s_base[tid] = (unsigned long long)atomicAdd(&out.ccount[s_q[tid]], s_cnt[tid]);
            } else {
                if (tid < QT && s_cnt[tid] > 0)
                    atomicAdd(&out.counts[s_q[tid]], (unsigned long long)s_cnt[tid]);
                if (MODE == MODE_IDS && tid == 0 && s_tile_n > 0)
                    s_tile_base = atomicAdd(out.res_n, (unsigned long long)s_tile_n);
            }
            __syncthreads();
            if (hitmask | bandmask) {
#pragma unroll
                for (int a = 0; a < 4; a++) {
                    const int i = ty * 4 + a;
#pragma unroll
                    for (int b = 0; b < 4; b++) {
                        const int bit = 1 << (a * 4 + b);
                        const int32_t pos = clid + r0 + tx * 4 + b;
                        if (hitmask & bit) {
                            const float d = tile_fin<M>(t, acc[a][b], s_qp[i], pos);
                            if (MODE == MODE_KNN) {
                                unsigned long long s = s_base[i] + slot_local[a][b];
                                if (s < (unsigned long long)out.cap)
                                    out.cand[(int64_t)s_q[i] * out.cap + s] =
                                        ((uint64_t)__float_as_uint(d) << 32) | (uint32_t)t.perm[pos];
                            } else if (MODE == MODE_IDS) {
                                unsigned long long s = s_tile_base + slot_local[a][b];
                                if ((int64_t)s < out.res_cap) {
                                    out.res_k[s] = ((uint64_t)(uint32_t)s_q[i] << 32) | (uint32_t)t.perm[pos];
                                    out.res_d[s] = d;
                                }
                            }
                        }
                        if (MODE != MODE_KNN && (bandmask & bit)) {
                            unsigned long long s = atomicAdd(out.band_n, 1ull);
                            if ((int64_t)s < out.band_cap) {
                                out.band_q[s] = s_q[i];
                                out.band_p[s] = pos;
                            }
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
    // verified-pair counter (one atomic per wave)
    for (int o = 32; o > 0; o >>= 1) nver += __shfl_xor(nver, o, 64);
    if ((tid & 63) == 0 && nver) atomicAdd(out.verified, nver);
}

// The L2 tile kernel on FP32 matrix cores: the same tiles as k_verify (up
// to 64 records of one parent node, that node's kept leaf rows in chunks of
// 64), with d^2 = |q|^2 + |x|^2 - 2 q.x and q.x from v_mfma_f32_16x16x4_f32.
// Wave w computes queries 32 (w & 1) .. +32 against rows 32 (w >> 1) .. +32
// as 2 x 2 16 x 16 MFMA tiles.  Each lane takes the 4 dimensions 4 (lane /
// 16) .. +3 of every 16-dimension group as four consecutive K steps (A and B
// use the same order, so the dot products are unchanged), one 16-byte LDS
// read per operand fragment.  Error: with c = fe.c1 (FP32 sums of at most
// dp + 4 terms), |d^2 - d_exact^2| <= E2 = c (|q| + |x|)^2, so a pair is
// decided as k_verify decides it with the one-sided bound min(sqrt(E2),
// E2 / d) (KeyErr's BF16-form band; the top-k re-ranks with the same model);
// counts only (range IDs keep the direct-difference kernel, whose reported
// distances are within 1e-5).
constexpr int MKP = DK + 4;   // LDS row pitch (floats) of the MFMA tiles: conflict-free 16-byte reads

// tile_slot[i]: the slot (parent node + 1) of tile i, from the tile offsets
__global__ void k_tile_slot(const int64_t *tile_off, int64_t nslots, int32_t *tile_slot)
{
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nslots;
         s += (int64_t)gridDim.x * blockDim.x)
        for (int64_t i = tile_off[s]; i < tile_off[s + 1]; i++) tile_slot[i] = (int32_t)s;
}

// One block per tile (the dispatcher balances tiles of uneven cost).
template <int MODE>
__global__ void __launch_bounds__(256, 4) k_verify_mfma(DevTree t, const int32_t *gq, const uint32_t *gm,
                                                     const int64_t *node_off, const int64_t *tile_off,
                                                     const int32_t *tile_slot, int64_t ntiles,
                                                     const int32_t *qpos, const float *bnd, BfErr fe,
                                                     VerifyOut out)
{
    __shared__ __attribute__((aligned(16))) float Qs[QT][MKP];
    __shared__ __attribute__((aligned(16))) float Xs[RT][MKP];
    __shared__ int32_t s_q[QT], s_qp[QT];
    __shared__ uint32_t s_m[QT];
    __shared__ float s_b[QT], s_qn[QT], s_xn[RT];
    __shared__ uint32_t s_union;
    __shared__ int32_t s_cnt[QT];
    __shared__ unsigned long long s_base[QT];

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    unsigned long long nver = 0;
    const int64_t tile = blockIdx.x;
    const int64_t slot = tile_slot[tile];
    const int64_t chunk = tile - tile_off[slot];
    const int64_t rstart = node_off[slot] + chunk * QT;
    const int rcount = (int)min((int64_t)QT, node_off[slot + 1] - rstart);
    const int32_t P = (int32_t)slot - 1;

    if (tid == 0) s_union = 0;
    __syncthreads();
    if (tid < QT) {
        if (tid < rcount) {
            const int32_t q = gq[rstart + tid];
            s_q[tid] = q;
            s_qp[tid] = qpos[q];
            s_qn[tid] = t.xn2[qpos[q]];
            s_m[tid] = gm[rstart + tid];
            s_b[tid] = bnd[q];
            atomicOr(&s_union, gm[rstart + tid]);
        } else {
            s_q[tid] = -1;
            s_qp[tid] = 0;
            s_qn[tid] = 0.0f;
            s_m[tid] = 0;
            s_b[tid] = -1.0f;
        }
    }
    __syncthreads();
    const uint32_t umask = s_union;
    const int lrow = tid >> 2, lk = (tid & 3) * 8;
    const int dp = t.dp;
    const int wq = 32 * (w & 1), wr = 32 * (w >> 1), kq = 4 * (lane >> 4), l16 = lane & 15;
    const int nchild = P < 0 ? 1 : t.T;
    const int32_t plid = P < 0 ? 0 : t.nlid[P];
    const int32_t pavg = P < 0 ? 0 : t.nsize[P] / t.T;

    for (int c = 0; c < nchild; c++) {
        if (!((umask >> c) & 1u)) continue;
        int cend = c;
        while (cend + 1 < nchild && ((umask >> (cend + 1)) & 1u)) cend++;
        int32_t clid, csz, elid, esz;
        child_range(t, P, c, clid, csz);
        child_range(t, P, cend, elid, esz);
        const int32_t run = elid + esz - clid;
        c = cend;
        // the thread's share of a slab (8 floats of one query row and of one
        // object row), loaded a slab ahead so that the loads of slab k + 1 are
        // in flight while slab k is multiplied; the last slab of a chunk loads
        // the next chunk's first slab
        const bool qv = lrow < rcount;
        const float *qrow = t.X + (int64_t)s_qp[lrow] * dp + lk;
        const float4 z4 = make_float4(0, 0, 0, 0);
        float4 p0 = (qv && lk < dp) ? *(const float4 *)qrow : z4;
        float4 p1 = (qv && lk + 4 < dp) ? *(const float4 *)(qrow + 4) : z4;
        float4 p2, p3;
        {
            const float *xrow = t.X + (int64_t)(clid + lrow) * dp + lk;
            p2 = (lrow < run && lk < dp) ? *(const float4 *)xrow : z4;
            p3 = (lrow < run && lk + 4 < dp) ? *(const float4 *)(xrow + 4) : z4;
        }
        for (int r0 = 0; r0 < run; r0 += RT) {
            f32x4 acc[2][2];
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
            const bool xv = r0 + lrow < run;
            const float *xrow = t.X + (int64_t)(clid + r0 + (xv ? lrow : 0)) * dp + lk;
            // the chunk's row norms, read once and kept in LDS for the epilogue
            const float xnv = (tid < RT && r0 + tid < run) ? t.xn2[clid + r0 + tid] : 0.0f;
            for (int k0 = 0; k0 < dp; k0 += DK) {
                // stage query rows and object rows (row-major, K contiguous;
                // a later chunk's first slab is already there)
                if (k0 > 0 || r0 == 0) {
                    *(float4 *)&Qs[lrow][lk] = p0;
                    *(float4 *)&Qs[lrow][lk + 4] = p1;
                    *(float4 *)&Xs[lrow][lk] = p2;
                    *(float4 *)&Xs[lrow][lk + 4] = p3;
                }
                __syncthreads();
                const int kn = k0 + DK;
                if (kn < dp) {
                    p0 = (qv && kn + lk < dp) ? *(const float4 *)(qrow + kn) : z4;
                    p1 = (qv && kn + lk + 4 < dp) ? *(const float4 *)(qrow + kn + 4) : z4;
                    p2 = (xv && kn + lk < dp) ? *(const float4 *)(xrow + kn) : z4;
                    p3 = (xv && kn + lk + 4 < dp) ? *(const float4 *)(xrow + kn + 4) : z4;
                } else if (r0 + RT < run) {
                    const bool nv = r0 + RT + lrow < run;
                    const float *nrow = t.X + (int64_t)(clid + r0 + RT + (nv ? lrow : 0)) * dp + lk;
                    p0 = (qv && lk < dp) ? *(const float4 *)qrow : z4;
                    p1 = (qv && lk + 4 < dp) ? *(const float4 *)(qrow + 4) : z4;
                    p2 = (nv && lk < dp) ? *(const float4 *)nrow : z4;
                    p3 = (nv && lk + 4 < dp) ? *(const float4 *)(nrow + 4) : z4;
                }
#pragma unroll
                for (int g = 0; g < DK; g += 16) {
                    float4 af[2], bf[2];
#pragma unroll
                    for (int a = 0; a < 2; a++) af[a] = *(const float4 *)&Qs[wq + 16 * a + l16][g + kq];
#pragma unroll
                    for (int b = 0; b < 2; b++) bf[b] = *(const float4 *)&Xs[wr + 16 * b + l16][g + kq];
#pragma unroll
                    for (int s = 0; s < 4; s++)
#pragma unroll
                        for (int a = 0; a < 2; a++)
#pragma unroll
                            for (int b = 0; b < 2; b++)
                                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32((&af[a].x)[s], (&bf[b].x)[s],
                                                                                 acc[a][b], 0, 0, 0);
                }
                __syncthreads();
            }
            if (r0 + RT < run) {   // the next chunk's first slab, staged before the epilogue
                *(float4 *)&Qs[lrow][lk] = p0;
                *(float4 *)&Qs[lrow][lk + 4] = p1;
                *(float4 *)&Xs[lrow][lk] = p2;
                *(float4 *)&Xs[lrow][lk + 4] = p3;
            }
            // epilogue: pair (query wq + 16a + kq + r, row wr + 16b + l16) in acc[a][b][r]
            if (tid < QT) s_cnt[tid] = 0;
            if (tid < RT) s_xn[tid] = xnv;
            __syncthreads();
            // a pair's slot within its query's share (< 64) in byte r of
            // slots[a][b]; its distance replaces its dot product in acc
            uint32_t slots[2][2] = {{0u, 0u}, {0u, 0u}};
            uint32_t hitmask = 0, bandmask = 0;
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < 2; b++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int i = wq + 16 * a + kq + r, j = wr + 16 * b + l16;
                        const int32_t pj = clid + r0 + j;
                        const int cj = P < 0 ? 0 : (pavg > 0 ? min((pj - plid) / pavg, t.T - 1) : t.T - 1);
                        const bool valid = i < rcount && r0 + j < run && ((s_m[i] >> cj) & 1u) &&
                                           !hidden(t, pj, s_q[i]);
                        if (!valid) {
                            acc[a][b][r] = 0.0f;
                            continue;
                        }
                        nver++;
                        const float qn = s_qn[i], xn = s_xn[j];
                        const bool self = s_qp[i] == pj;
                        const float d = self ? 0.0f : sqrtf(fmaxf(qn + xn - 2.0f * acc[a][b][r], 0.0f));
                        const float sq = sqrtf(qn) + sqrtf(xn);
                        const float E2 = fe.c1 * sq * sq;
                        const float e = fminf(sqrtf(E2), d > 0.0f ? E2 / d : INFINITY) * 1.0001f + 2.4e-7f * d;
                        acc[a][b][r] = d;
                        const int bit = 1 << (a * 8 + b * 4 + r);
                        const float bq = s_b[i];
                        if (MODE == MODE_KNN) {
                            if (self || d - e <= bq) {
                                hitmask |= bit;
                                slots[a][b] |= (uint32_t)atomicAdd(&s_cnt[i], 1) << (8 * r);
                            }
                        } else {
                            if (!self && d + e > bq) {
                                if (d - e <= bq) bandmask |= bit;
                            } else {
                                hitmask |= bit;
                                atomicAdd(&s_cnt[i], 1);
                            }
                        }
                    }
            __syncthreads();
            if (MODE == MODE_KNN) {
                if (tid < QT && s_cnt[tid] > 0)
                    s_base[tid] = (unsigned long long)atomicAdd(&out.ccount[s_q[tid]], s_cnt[tid]);
            } else {
                if (tid < QT && s_cnt[tid] > 0) atomicAdd(&out.counts[s_q[tid]], (unsigned long long)s_cnt[tid]);
            }
            __syncthreads();
            if (hitmask | bandmask) {
#pragma unroll
                for (int a = 0; a < 2; a++)
#pragma unroll
                    for (int b = 0; b < 2; b++)
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                            const int bit = 1 << (a * 8 + b * 4 + r);
                            const int i = wq + 16 * a + kq + r;
                            const int32_t pos = clid + r0 + wr + 16 * b + l16;
                            if (MODE == MODE_KNN && (hitmask & bit)) {
                                const unsigned long long s = s_base[i] + ((slots[a][b] >> (8 * r)) & 0xffu);
                                if (s < (unsigned long long)out.cap)
                                    out.cand[(int64_t)s_q[i] * out.cap + s] =
                                        ((uint64_t)__float_as_uint(acc[a][b][r]) << 32) | (uint32_t)t.perm[pos];
                            }
                            if (MODE != MODE_KNN && (bandmask & bit)) {
                                const unsigned long long s = atomicAdd(out.band_n, 1ull);
                                if ((int64_t)s < out.band_cap) {
                                    out.band_q[s] = s_q[i];
                                    out.band_p[s] = pos;
                                }
                            }
                        }
            }
            __syncthreads();
        }
    }
    for (int o = 32; o > 0; o >>= 1) nver += __shfl_xor(nver, o, 64);
    if (lane == 0 && nver) atomicAdd(out.verified, nver);
}

// Edit distance: one wave per record, lanes over the objects of its leaves;
// exact integer distances, so there is no band.
// Myers / Hyyro bit-parallel edit distance with the pattern's match masks in
// peq (bit i of peq[c] set iff pattern[i] == c; pattern length 1 <= m <= the
// word width) over the text bytes held in tw (n of them, n <= 64): one
// column of vertical +1/-1 deltas per word (bit i of Pv / Mv: D[i+1][j] -
// D[i][j] = +1 / -1), O(n) word operations.  Row 0 is D[0][j] = j (the +1
// shifted in at the bottom of every horizontal delta), so the distance is
// D[m][n] = n + |Pv| - |Mv| over the m pattern bits of the last column;
// bits above m never reach bits below it (carries and shifts go upward).
template <typename W>
__device__ __forceinline__ int myers_peq(const uint64_t *peq, int m, const uint32_t (&tw)[16], int n)
{
    W Pv = ~W(0), Mv = W(0);
#pragma unroll
    for (int w = 0; w < 16; w++) {
        if (4 * w >= n) break;
        const uint32_t word = tw[w];
#pragma unroll
        for (int b = 0; b < 4; b++) {
            if (4 * w + b >= n) break;
            const W Eq = (W)peq[(word >> (8 * b)) & 255u];
            const W Xv = Eq | Mv;
            const W Xh = (((Eq & Pv) + Pv) ^ Pv) | Eq;
            const W Ph = (Mv | ~(Xh | Pv)) << 1 | W(1);
            const W Mh = (Pv & Xh) << 1;
            Pv = Mh | ~(Xv | Ph);
            Mv = Ph & Xv;
        }
    }
    const W mask = m >= (int)(8 * sizeof(W)) ? ~W(0) : (W(1) << m) - W(1);
    return n + __popcll((uint64_t)(Pv & mask)) - __popcll((uint64_t)(Mv & mask));
}

// Edit-distance verification of the leaf records (dataProcessRnn /
// dataProcessKnn with the edit metric, search_v2.cuh:388-414, 319-544): one
// wave per record (query, parent, kept-leaf mask).  The objects of all kept
// leaves of the record are laid out over the wave's lanes (a prefix over the
// kept children), so a wave verifies 64 of them at a time whatever the leaf
// size.  The query is the pattern of the bit-parallel algorithm when it has
// at most 64 bytes: its match masks are built once per record in the wave's
// LDS table (one atomic OR per pattern byte), each text row is read as
// 16-byte words into registers, and a text byte then costs one LDS read and
// about a dozen word operations (32-bit words for queries of at most 32
// bytes); the distance comes from the last column's deltas.
// Longer queries use gts_edit_distance.  Results are aggregated per wave (one
// atomic per 64 objects).
template <int MODE>
__global__ void __launch_bounds__(256) k_verify_edit(DevTree t, const int32_t *gq,
                                                     const int32_t *gn, const uint32_t *gm,
                                                     int64_t nrec, const int32_t *qpos,
                                                     const float *bnd, VerifyOut out)
{
    __shared__ uint64_t s_peq[4][256];
    __shared__ int32_t s_cl[4][33], s_co[4][33];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint64_t *peq = s_peq[wv];
    int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int nv = t.sstride / 16;   // 16-byte words per string row (t.sstride is a multiple of 16)
    const uint64_t lower = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    unsigned long long nver = 0;
    for (int64_t r = w; r < nrec; r += nw) {
        const int32_t q = gq[r], P = gn[r];
        const uint32_t m = gm[r];
        const int32_t qp = qpos[q];
        const int qlen = t.slen[qp];
        const uint8_t *qs = t.S + (int64_t)qp * t.sstride;
        const float b = bnd[q];
        const bool pat = qlen >= 1 && qlen <= 64;
        if (pat) {
            // (a wave's LDS instructions execute in order: the previous
            // record's reads come before this zeroing, the zeroing before
            // the ORs)
            for (int c = lane; c < 256; c += 64) peq[c] = 0ull;
            if (lane < qlen) atomicOr((unsigned long long *)&peq[qs[lane]], 1ull << lane);
        }
        // the kept children with objects, compacted; prefix of their sizes
        const int nchild = P < 0 ? 1 : t.T;
        int32_t clid = 0, csz = 0;
        if (lane < nchild && ((m >> lane) & 1u)) child_range(t, P, lane, clid, csz);
        const bool has = csz > 0;
        const uint64_t hbal = __builtin_amdgcn_ballot_w64(has);
        const int rank = (int)__popcll(hbal & lower), nk = (int)__popcll(hbal);
        int incl = csz;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(incl, o, 64);
            if (lane >= o) incl += u;
        }
        const int total = __shfl(incl, 63, 64);
        if (has) {
            s_cl[wv][rank] = clid;
            s_co[wv][rank] = incl - csz;
        }
        int c = 0;
        for (int j0 = 0; j0 < total; j0 += 64) {
            const int j = j0 + lane;
            bool hit = false;
            float d = 0.0f;
            int32_t pos = 0;
            if (j < total) {
                while (c + 1 < nk && s_co[wv][c + 1] <= j) c++;
                pos = s_cl[wv][c] + (j - s_co[wv][c]);
                if (!hidden(t, pos, q)) {
                    const int xl = t.slen[pos];
                    nver++;
                    if ((float)abs(xl - qlen) <= b) {
                        const uint8_t *xs = t.S + (int64_t)pos * t.sstride;
                        int de;
                        if (pat && xl <= 64) {
                            uint32_t tw[16];
#pragma unroll
                            for (int v = 0; v < 4; v++) {
                                const uint4 u4 = v < nv ? ((const uint4 *)xs)[v] : make_uint4(0, 0, 0, 0);
                                tw[4 * v] = u4.x;
                                tw[4 * v + 1] = u4.y;
                                tw[4 * v + 2] = u4.z;
                                tw[4 * v + 3] = u4.w;
                            }
                            de = qlen <= 32 ? myers_peq<uint32_t>(peq, qlen, tw, xl)
                                            : myers_peq<uint64_t>(peq, qlen, tw, xl);
                        } else {
                            de = gts_edit_distance(qs, qlen, xs, xl);
                        }
                        d = (float)de;
                        hit = d <= b;
                    }
                }
            }
            const uint64_t hm = __builtin_amdgcn_ballot_w64(hit);
            if (!hm) continue;
            const int cnt = (int)__popcll(hm), me = (int)__popcll(hm & lower);
            const int first = __builtin_ctzll(hm);
            if (MODE == MODE_KNN) {
                int base = 0;
                if (lane == first) base = atomicAdd(&out.ccount[q], cnt);
                base = __shfl(base, first, 64);
                const int s = base + me;
                if (hit && s < out.cap)
                    out.cand[(int64_t)q * out.cap + s] = ((uint64_t)__float_as_uint(d) << 32) | (uint32_t)t.perm[pos];
            } else {
                if (lane == first) atomicAdd(&out.counts[q], (unsigned long long)cnt);
                if (MODE == MODE_IDS) {
                    unsigned long long base = 0;
                    if (lane == first) base = atomicAdd(out.res_n, (unsigned long long)cnt);
                    base = __shfl(base, first, 64);
                    const unsigned long long s = base + (unsigned long long)me;
                    if (hit && (int64_t)s < out.res_cap) {
                        out.res_k[s] = ((uint64_t)(uint32_t)q << 32) | (uint32_t)t.perm[pos];
                        out.res_d[s] = d;
                    }
                }
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) nver += __shfl_xor(nver, o, 64);
    if (lane == 0 && nver) atomicAdd(out.verified, nver);
}

// FP64 re-decision of range pairs inside the FP32 band (one wave per pair).
template <int M, int MODE>
__global__ void __launch_bounds__(256) k_band_range(DevTree t, const int32_t *band_q,
                                                    const int32_t *band_p, int64_t nb,
                                                    const int32_t *qpos, const float *bnd,
                                                    VerifyOut out)
{
    const int lane = threadIdx.x & 63;
    int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t i = w; i < nb; i += nw) {
        const int32_t q = band_q[i], pos = band_p[i];
        const int32_t qp = qpos[q];
        double d = group_dist64<M, 64>(t.X + (int64_t)qp * t.dp, t.X + (int64_t)pos * t.dp, t.dim, lane);
        if (qp == pos) d = 0.0;
        if (lane == 0 && d <= (double)bnd[q]) {
            atomicAdd(&out.counts[q], 1ull);
            if (MODE == MODE_IDS) {
                unsigned long long s = atomicAdd(out.res_n, 1ull);
                if ((int64_t)s < out.res_cap) {
                    out.res_k[s] = ((uint64_t)(uint32_t)q << 32) | (uint32_t)t.perm[pos];
                    out.res_d[s] = (float)d;
                }
            }
        }
    }
}

// -------------------------------------------------------------- top-k
constexpr int TOPK_CAP = 4096;  // candidates per query held in LDS
constexpr int FP32_MFMA_MIN_DIM = 32;   // default: FP32-form L2 tiles on the matrix cores from this dim
constexpr int TOPK_DEFERRED = 3; // status: left by the wave-level kernel for the block-level one
constexpr int WAVE_K = 64;       // largest k of the wave-level kernel
constexpr int WBAND = 256;       // FP64 band capacity per query of the wave-level kernel

__device__ void bitonic_u64(uint64_t *a, int n)
{
    for (int k = 2; k <= n; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < n; i += blockDim.x) {
                int ixj = i ^ j;
                if (ixj > i) {
                    uint64_t x = a[i], y = a[ixj];
                    bool up = (i & k) == 0;
                    if ((x > y) == up) {
                        a[i] = y;
                        a[ixj] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// pairs (key, id) ordered lexicographically
__device__ void bitonic_pair(uint64_t *key, uint32_t *id, int n)
{
    for (int k = 2; k <= n; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < n; i += blockDim.x) {
                int ixj = i ^ j;
                if (ixj > i) {
                    uint64_t x = key[i], y = key[ixj];
                    uint32_t xi = id[i], yi = id[ixj];
                    bool gt = x > y || (x == y && xi > yi);
                    bool up = (i & k) == 0;
                    if (gt == up) {
                        key[i] = y; key[ixj] = x;
                        id[i] = yi; id[ixj] = xi;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Exact k smallest by (distance, ID) of each query's candidates: sort by the
// FP32 key, take every candidate whose lower error bound reaches the k-th
// upper bound, re-rank those by FP64 distance, emit.  Status per query:
// 0 done, 1 candidate overflow (bound tightened, rerun), 2 the k-th exact
// distance exceeds the radius that filtered the candidates (rerun wider).
// Error bound of a candidate's FP32 key: the FP32 model rel*d + abs, or for
// keys from the BF16 form min(sqrt(E2), E2/d) with the query's worst-case E2.
struct KeyErr {
    ErrModel em;
    int bf;
    float E2;
    __device__ __forceinline__ float operator()(float d) const
    {
        if (!bf) return em.rel * d + em.abs;
        return fminf(sqrtf(E2), d > 0.0f ? E2 / d : INFINITY);
    }
    // One-sided bound |d_exact - d| of a key's distance, as the band and the
    // bound updates use it.  FP32 keys: twice the model.  BF16 keys: with
    // |D - d_exact^2| <= E2 for the key's d^2 = D (E2 includes the FP32
    // rounding of D), |sqrt(D) - d_exact| <= min(sqrt(E2), E2 / sqrt(D)); the
    // 10^-4 relative and 4-ulp terms cover the rounding of this bound, of the
    // key's square root and of the comparisons.
    __device__ __forceinline__ float band(float d) const
    {
        if (!bf) return 2.0f * (em.rel * d + em.abs);
        return fminf(sqrtf(E2), d > 0.0f ? E2 / d : INFINITY) * 1.0001f + 2.4e-7f * d;
    }
};

// A query's final status; a nonzero one also joins the pass's failure list
// (fail[], count in nfail[0]) when the caller collects it.
__device__ __forceinline__ void final_status(int32_t *status, int64_t q, int st, int32_t *fail,
                                             unsigned long long *nfail)
{
    status[q] = st;
    if (st != 0 && fail) fail[atomicAdd(nfail, 1ull)] = (int32_t)q;
}

template <int M>
__global__ void __launch_bounds__(256) k_topk(DevTree t, const int32_t *inv, const uint64_t *cand,
                                              const int32_t *ccount, int cap, const int32_t *qpos,
                                              const int32_t *kq, float *bnd, int check_radius,
                                              ErrModel em, int bf, BfErr be, const float *xn2,
                                              float xn2_max, const int32_t *orig, int kstride,
                                              int32_t *ids, float *dists, float *kth,
                                              int32_t *status, int only_deferred, int32_t *fail,
                                              unsigned long long *ctr, unsigned long long *ncand)
{
    __shared__ uint64_t keys[TOPK_CAP];
    __shared__ uint64_t d64[TOPK_CAP];
    __shared__ uint64_t s_top[4][64];
    __shared__ uint64_t sh_kkey;
    __shared__ int sh_ne;
    const int q = blockIdx.x;
    const int tid = threadIdx.x;
    // second pass after the wave-level kernel: only the queries it deferred
    if (only_deferred && status[q] != TOPK_DEFERRED) return;
    if (orig[q] < 0) {   // an inert padding query of a rerun batch
        if (tid == 0) status[q] = 0;
        return;
    }
    const int total = ccount[q];
    if (ncand && !only_deferred && tid == 0 && total) atomicAdd(ncand, (unsigned long long)total);
    const int cnt = min(total, cap);
    const int k = kq[q];
    const int orow = orig[q];
    KeyErr ke;
    ke.em = em;
    ke.bf = bf;
    ke.E2 = 0.0f;
    if (bf) {
        float qn = xn2[qpos[q]];
        // the key's d^2 is rounded in FP32 too: 8 ulps of |q|^2 + |x|^2
        ke.E2 = (be.c1 + 4.8e-7f) * (qn + xn2_max) + 2.0f * be.c2 * sqrtf(qn * xn2_max) + 2.0f * sqrtf(qn) * be.ex;
    }
    const int kk = min(k, cnt);
    const uint64_t *cq = cand + (int64_t)q * cap;
    if (tid == 0) sh_ne = 0;
    uint64_t kkey = ~0ull;   // the kk-th smallest stored key
    if (kk <= WAVE_K) {
        // selection: each wave keeps the kk smallest keys of its slices in a
        // running top-64 (slices with no key below its kk-th are skipped),
        // wave 0 merges the four lists; the band is then collected unsorted
        const int lane = tid & 63, wv = tid >> 6;
        uint64_t best = ~0ull;
        if (kk > 0)
            for (int b0 = wv * 64; b0 < cnt; b0 += (int)blockDim.x) {
                uint64_t v = b0 + lane < cnt ? cq[b0 + lane] : ~0ull;
                if (!__builtin_amdgcn_ballot_w64(v < lane_key64(best, kk - 1))) continue;
                v = wave_sort64(v, lane);
                best = wave_clean64(min(best, lane_read64(v, 63 - lane)), lane);
            }
        s_top[wv][lane] = best;
        __syncthreads();
        if (wv == 0) {
            uint64_t m = s_top[0][lane];
            for (int w2 = 1; w2 < (int)(blockDim.x >> 6); w2++) m = wave_merge64(m, s_top[w2][lane], lane);
            const uint64_t kth_key = lane_read64(m, max(kk, 1) - 1);
            if (lane == 0) sh_kkey = kth_key;
        }
        __syncthreads();
        kkey = sh_kkey;
    } else {
        int np = 1;
        while (np < cnt) np <<= 1;
        for (int i = tid; i < np; i += blockDim.x) keys[i] = i < cnt ? cq[i] : ~0ull;
        __syncthreads();
        bitonic_u64(keys, np);
        kkey = keys[kk - 1];
    }
    if (total > cap) {
        // overflow: the k-th of the stored candidates is a valid bound
        if (tid == 0) {
            if (kk == k) {
                float kthf = __uint_as_float((uint32_t)(kkey >> 32));
                float up = kthf + ke.band(kthf);
                if (up < bnd[q]) bnd[q] = up;
            }
            final_status(status, q, 1, fail, ctr);
        }
        return;
    }
    // band: candidates whose lower bound does not exceed the k-th upper bound
    // (a prefix of the sorted keys, since d - 2 e(d) grows with d)
    if (kk > 0) {
        const float tk = __uint_as_float((uint32_t)(kkey >> 32));
        const float U = tk + ke.band(tk);
        if (kk <= WAVE_K) {
            __syncthreads();   // keys[] is free: collect the band there, unsorted
            for (int i = tid; i < cnt; i += blockDim.x) {
                const uint64_t v = cq[i];
                const float d = __uint_as_float((uint32_t)(v >> 32));
                if (d - ke.band(d) <= U) keys[atomicAdd(&sh_ne, 1)] = v;
            }
        } else {
            int local = 0;
            for (int i = tid; i < cnt; i += blockDim.x) {
                float d = __uint_as_float((uint32_t)(keys[i] >> 32));
                if (d - ke.band(d) <= U) local = max(local, i + 1);
            }
            atomicMax(&sh_ne, local);
        }
    }
    __syncthreads();
    const int ne = max(sh_ne, kk);
    // FP64 re-ranking of the first ne candidates: eight lanes per candidate,
    // 16-byte row loads, blockDim / 8 candidates at a time
    const int32_t qp = qpos[q];
    if (M == GTS_METRIC_EDIT) {
        for (int i = tid; i < ne; i += blockDim.x)
            d64[i] = (uint64_t)__double_as_longlong((double)__uint_as_float((uint32_t)(keys[i] >> 32)));
    } else {
        constexpr int G = 8;
        const int grp = tid / G, gl = tid % G;
        for (int i0 = 0; i0 < ne; i0 += blockDim.x / G) {
            const int i = i0 + grp;
            const bool ok = i < ne;
            const int32_t pos = ok ? inv[(uint32_t)keys[i]] : qp;
            double d = group_dist64_v4<M == GTS_METRIC_EDIT ? GTS_METRIC_L2 : M, G>(
                t.X + (int64_t)qp * t.dp, t.X + (int64_t)pos * t.dp, t.dp, gl);
            if (pos == qp) d = 0.0;
            if (ok && gl == 0) d64[i] = (uint64_t)__double_as_longlong(d);
        }
    }
    __syncthreads();
    // sort (d64, id) over the first ne entries; ids live in the low key words
    uint32_t *idv = (uint32_t *)keys;  // reuse: ids compacted into the first ne words
    __syncthreads();
    {
        // compact ids: keys[i] low word -> idv[i]; do it via registers to avoid aliasing
        uint32_t tmp[TOPK_CAP / 256];
        for (int r = 0; r < TOPK_CAP / 256; r++) {
            int i = tid + r * 256;
            tmp[r] = i < ne ? (uint32_t)keys[i] : 0xffffffffu;
        }
        __syncthreads();
        int np2 = 1;
        while (np2 < ne) np2 <<= 1;
        for (int r = 0; r < TOPK_CAP / 256; r++) {
            int i = tid + r * 256;
            if (i < np2) idv[i] = i < ne ? tmp[r] : 0xffffffffu;
            if (i >= ne && i < np2) d64[i] = ~0ull;
        }
        __syncthreads();
        bitonic_pair(d64, idv, np2);
    }
    for (int i = tid; i < kstride; i += blockDim.x) {
        if (i < kk) {
            ids[(int64_t)orow * kstride + i] = (int32_t)idv[i];
            dists[(int64_t)orow * kstride + i] = (float)__longlong_as_double((long long)d64[i]);
        } else {
            ids[(int64_t)orow * kstride + i] = -1;
            dists[(int64_t)orow * kstride + i] = INFINITY;
        }
    }
    if (tid == 0) {
        int st = 0;
        if (check_radius && !isinf(bnd[q])) {
            double dk = kk > 0 ? __longlong_as_double((long long)d64[kk - 1]) : INFINITY;
            if (kk < k || dk > (double)bnd[q]) st = 2;
        }
        // the k-th distance does not exist when fewer than k objects qualify
        kth[orow] = kk == k ? (float)__longlong_as_double((long long)d64[kk - 1]) : INFINITY;
        final_status(status, q, st, fail, ctr);
    }
}

// ---------------------------------------------------- wave-level top-k
//
// One wave per query (4 per block), k <= 64.  The wave keeps the 64 smallest
// FP32 keys seen so far, one per lane in ascending lane order.  Each chunk of
// 64 candidates is sorted across the lanes by a bitonic network and merged:
// min(best[i], chunk[63 - i]) is bitonic, one cleaning pass sorts it.  The
// lane exchanges are DPP quad permutations (xor 1, 2), ds_swizzle in bit-mask
// mode (xor 4, 8, 16) and ds_bpermute (xor 32, across the two 32-lane halves).
// A second pass collects the error band around the k-th key, the band is
// re-ranked by FP64 distance and (distance, ID) ranks place the first k.  A
// query whose band exceeds WBAND is deferred to the block-level kernel.
// The wave's steps are chains of dependent loads (list, band IDs, rows), so
// residency is what hides them: at most 80 registers, 6 waves per SIMD
// (angular, with three sums per lane, keeps 4).

template <int M>
__global__ void __launch_bounds__(256, M == GTS_METRIC_ANGULAR ? 4 : 6) k_topk_wave(DevTree t, const int32_t *inv, const uint64_t *cand,
                                                   const int32_t *ccount, int cap, const int32_t *qpos,
                                                   const int32_t *kq, float *bnd, int check_radius,
                                                   ErrModel em, int bf, BfErr be, const float *xn2,
                                                   float xn2_max, const int32_t *orig, int kstride,
                                                   int32_t *ids, float *dists, float *kth, int32_t *status,
                                                   int64_t nq, int32_t *fail, unsigned long long *ctr,
                                                   unsigned long long *ncand)
{
    __shared__ uint32_t s_id[4][WBAND];
    __shared__ double s_d[4][WBAND];
    __shared__ uint64_t s_sel[4][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t q = (int64_t)blockIdx.x * 4 + wv;
    if (q >= nq) return;   // the whole wave: no block barrier below
    if (orig[q] < 0) {     // an inert padding query of a rerun batch
        if (lane == 0) status[q] = 0;
        return;
    }
    const int total = ccount[q];
    if (ncand && lane == 0 && total) atomicAdd(ncand, (unsigned long long)total);
    const int cnt = min(total, cap);
    const int k = kq[q];
    const int orow = orig[q];
    KeyErr ke;
    ke.em = em;
    ke.bf = bf;
    ke.E2 = 0.0f;
    if (bf) {
        const float qn = xn2[qpos[q]];
        // the key's d^2 is rounded in FP32 too: 8 ulps of |q|^2 + |x|^2
        ke.E2 = (be.c1 + 4.8e-7f) * (qn + xn2_max) + 2.0f * be.c2 * sqrtf(qn * xn2_max) + 2.0f * sqrtf(qn) * be.ex;
    }
    const uint64_t *cq = cand + q * (int64_t)cap;
    const int kk = min(k, cnt);
    // The kk-th smallest of the 64 lane minima bounds the kk-th smallest key
    // (the kk smallest lane minima are kk distinct candidates), so the band
    // test below, taken at that bound, keeps every candidate of the band and
    // the kk smallest keys: those are gathered into one 64-key slice (second
    // pass) and sorted once.  More than 64 of them fall back to a running
    // top-64 over the list and a separate band pass.
    // Lists of at most 64 RK keys are read once, into registers (every load
    // in flight together); longer ones are streamed twice.
    constexpr int RK = 16;
    const bool inreg = cnt <= 64 * RK;
    uint64_t rk[RK];
    uint64_t mn = ~0ull;
    if (inreg) {
#pragma unroll
        for (int u = 0; u < RK; u++) rk[u] = 64 * u + lane < cnt ? cq[64 * u + lane] : ~0ull;
#pragma unroll
        for (int u = 0; u < RK; u++) mn = min(mn, rk[u]);
    } else {
        for (int b0 = 0; b0 < cnt; b0 += 256) {
            uint64_t vs[4];   // four slices' loads in flight
#pragma unroll
            for (int u = 0; u < 4; u++) vs[u] = b0 + 64 * u + lane < cnt ? cq[b0 + 64 * u + lane] : ~0ull;
#pragma unroll
            for (int u = 0; u < 4; u++) mn = min(mn, vs[u]);
        }
    }
    // U(t) = t + band(t) is the band's upper limit for a k-th distance t;
    // its largest value over t <= Tb: at Tb or, for the BF16 model (band(t)
    // ~ min(sqrt(E2), E2 / t)), at t = sqrt(E2)
    auto band_limit = [&](float Tb) {
        float u = Tb + ke.band(Tb);
        if (bf) {
            const float t = fminf(Tb, sqrtf(ke.E2));
            u = fmaxf(u, t + ke.band(t));
        }
        return u;
    };
    int nsel = 65;
    uint64_t best = ~0ull;
    if (kk > 0) {
        // only the distance of the kk-th smallest lane minimum is used: sort
        // the minima's distance bits (non-negative floats order as their bits)
        const uint32_t T = (uint32_t)__shfl((int)wave_sort32((uint32_t)(mn >> 32), lane), kk - 1, 64);
        const float Ub = band_limit(__uint_as_float(T)) * 1.000001f;
        nsel = 0;
        auto gather = [&](uint64_t v, int i) {
            const float d = __uint_as_float((uint32_t)(v >> 32));
            const bool in = i < cnt && d - ke.band(d) <= Ub;
            const uint64_t m = __builtin_amdgcn_ballot_w64(in);
            const int pos = nsel + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                             __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
            if (in && pos < 64) s_sel[wv][pos] = v;
            nsel += __popcll(m);
        };
        if (inreg) {
#pragma unroll
            for (int u = 0; u < RK; u++)
                if (64 * u < cnt) gather(rk[u], 64 * u + lane);
        } else {
            for (int b0 = 0; b0 < cnt; b0 += 256) {
                uint64_t vs[4];
#pragma unroll
                for (int u = 0; u < 4; u++) vs[u] = b0 + 64 * u + lane < cnt ? cq[b0 + 64 * u + lane] : ~0ull;
#pragma unroll
                for (int u = 0; u < 4; u++) gather(vs[u], b0 + 64 * u + lane);
            }
        }
    }
    int nb = 0;
    if (nsel <= 64) {
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        best = wave_sort64(lane < nsel ? s_sel[wv][lane] : ~0ull, lane);
    } else {
        for (int b0 = 0; b0 < cnt; b0 += 256) {
            uint64_t vs[4];   // four slices' loads in flight
#pragma unroll
            for (int u = 0; u < 4; u++) vs[u] = b0 + 64 * u + lane < cnt ? cq[b0 + 64 * u + lane] : ~0ull;
#pragma unroll
            for (int u = 0; u < 4; u++) {
                uint64_t v = vs[u];
                // a slice with no key below the current kk-th smallest changes nothing
                if (kk > 0 && !__builtin_amdgcn_ballot_w64(v < lane_key64(best, kk - 1))) continue;
                v = wave_sort64(v, lane);
                best = wave_clean64(min(best, lane_read64(v, 63 - lane)), lane);
            }
        }
    }
    // the kk-th smallest key, read with every lane active (cross-lane reads
    // from inactive lanes are undefined)
    const uint64_t kkey = lane_read64(best, max(kk, 1) - 1);
    if (total > cap) {
        // overflow: the k-th of the stored candidates is a valid bound
        if (lane == 0) {
            if (kk == k) {
                const float kthf = __uint_as_float((uint32_t)(kkey >> 32));
                const float up = kthf + ke.band(kthf);
                if (up < bnd[q]) bnd[q] = up;
            }
            final_status(status, q, 1, fail, ctr);
        }
        return;
    }
    // band: candidates whose lower bound does not exceed the k-th upper bound
    if (kk > 0) {
        const float tk = __uint_as_float((uint32_t)(kkey >> 32));
        const float U = tk + ke.band(tk);
        if (nsel <= 64) {
            // the gathered slice holds the band: keep its members, in place
            const float d = __uint_as_float((uint32_t)(best >> 32));
            const bool in = lane < nsel && d - ke.band(d) <= U;
            const uint64_t m = __builtin_amdgcn_ballot_w64(in);
            const int pos = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
            if (in) {
                s_id[wv][pos] = (uint32_t)best;
                s_d[wv][pos] = (double)d;
            }
            nb = __popcll(m);
        } else {
            for (int b0 = 0; b0 < cnt; b0 += 256) {
                uint64_t vs[4];   // four slices' loads in flight
#pragma unroll
                for (int u = 0; u < 4; u++) vs[u] = b0 + 64 * u + lane < cnt ? cq[b0 + 64 * u + lane] : ~0ull;
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const bool valid = b0 + 64 * u + lane < cnt;
                    const uint64_t v = vs[u];
                    const float d = __uint_as_float((uint32_t)(v >> 32));
                    const bool in = valid && d - ke.band(d) <= U;
                    const uint64_t m = __builtin_amdgcn_ballot_w64(in);
                    const int pos = nb + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
                    if (in && pos < WBAND) {
                        s_id[wv][pos] = (uint32_t)v;
                        s_d[wv][pos] = (double)d;   // edit: the exact distance; others: replaced below
                    }
                    nb += __popcll(m);
                }
            }
        }
    }
    if (nb > WBAND) {
        if (lane == 0) status[q] = TOPK_DEFERRED;
        return;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    // FP64 re-ranking of the band: eight candidates at a time, eight lanes
    // each (their row loads in flight together; the next group's row
    // positions loaded while this group's distances are formed)
    const int32_t qp = qpos[q];
    auto fp64 = [&](int a, int b) {
        constexpr int G = 8;
        const int grp = lane / G, gl = lane % G;
        int32_t pos = a + grp < b ? inv[s_id[wv][a + grp]] : qp;
        for (int i0 = a; i0 < b; i0 += 64 / G) {
            const int i = i0 + grp, in = i + 64 / G;
            const bool ok = i < b;
            const int32_t pnext = in < b ? inv[s_id[wv][in]] : qp;
            double d = group_dist64_v4<M == GTS_METRIC_EDIT ? GTS_METRIC_L2 : M, G>(
                t.X + (int64_t)qp * t.dp, t.X + (int64_t)pos * t.dp, t.dp, gl);
            if (pos == qp) d = 0.0;
            if (ok && gl == 0) s_d[wv][i] = d;
            pos = pnext;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    };
    if (M != GTS_METRIC_EDIT) {
        if (nsel <= 64 && kk > 0 && nb > kk) {
            // The band is in ascending key order, its first kk entries the kk
            // smallest keys: their FP64 distances first.  T, their largest,
            // bounds the kk-th smallest distance (kk distinct rows lie at or
            // below it), so a later entry whose lower bound exceeds T has kk
            // rows strictly closer and is dropped before its row is read.
            fp64(0, kk);
            double T = 0.0;
            for (int i = lane; i < kk; i += 64) T = fmax(T, s_d[wv][i]);
            for (int o = 32; o > 0; o >>= 1) T = fmax(T, __shfl_xor(T, o, 64));
            const int i = kk + lane;   // nb <= 64
            const bool valid = i < nb;
            const float d = valid ? (float)s_d[wv][i] : 0.0f;   // the key's distance (not yet replaced)
            const uint32_t id = valid ? s_id[wv][i] : 0u;
            const bool keep = valid && (double)(d - ke.band(d)) <= T * (1.0 + 1e-6);
            const uint64_t m = __builtin_amdgcn_ballot_w64(keep);
            const int pos = kk + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
            __builtin_amdgcn_wave_barrier();   // every entry read before the compaction writes
            if (keep) {
                s_id[wv][pos] = id;
                s_d[wv][pos] = (double)d;
            }
            nb = kk + __popcll(m);
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            fp64(kk, nb);
        } else {
            fp64(0, nb);
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    // rank by (FP64 distance, ID) and place the first kk
    for (int i = lane; i < nb; i += 64) {
        const double di = s_d[wv][i];
        const uint32_t ii = s_id[wv][i];
        int rank = 0;
        for (int j = 0; j < nb; j++) {
            const double dj = s_d[wv][j];
            rank += dj < di || (dj == di && s_id[wv][j] < ii);
        }
        if (rank < kk) {
            ids[(int64_t)orow * kstride + rank] = (int32_t)ii;
            dists[(int64_t)orow * kstride + rank] = (float)di;
        }
        if (rank == kk - 1) {
            int st = 0;
            if (check_radius && !isinf(bnd[q]) && (kk < k || di > (double)bnd[q])) st = 2;
            kth[orow] = kk == k ? (float)di : INFINITY;
            final_status(status, q, st, fail, ctr);
        }
    }
    for (int i = kk + lane; i < kstride; i += 64) {
        ids[(int64_t)orow * kstride + i] = -1;
        dists[(int64_t)orow * kstride + i] = INFINITY;
    }
    if (kk == 0 && lane == 0) {
        kth[orow] = INFINITY;
        final_status(status, q, (check_radius && !isinf(bnd[q])) ? 2 : 0, fail, ctr);
    }
}

}  // namespace

// ---------------------------------------------------------------- host side

// Times the verification launches of one call with HIP events on the index
// stream (the stream the kernels run on); the elapsed time is added to the
// call's statistics once the events have completed.
struct EventTimer {
    hipStream_t st;
    gts_stats *s;
    hipEvent_t a, b;
    bool on;
    EventTimer(hipStream_t st_, gts_stats *s_, bool on_) : st(st_), s(s_), on(on_)
    {
        if (!on) return;
        GTS_HIP(hipEventCreate(&a));
        GTS_HIP(hipEventCreate(&b));
        GTS_HIP(hipEventRecord(a, st));
    }
    ~EventTimer()
    {
        if (!on) return;
        if (hipEventRecord(b, st) == hipSuccess && hipEventSynchronize(b) == hipSuccess) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, a, b) == hipSuccess) {
                s->ms_verify += ms;
                s->verify_launches++;
            }
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
};

void Search::ensure_records(int64_t need)
{
    if ((int64_t)idx->rec_q.cap < need) {
        // keep existing records when growing
        int64_t keep = nrec;
        int32_t *oq = nullptr, *on = nullptr;
        uint32_t *om = nullptr;
        if (keep > 0) {
            GTS_HIP(hipMallocAsync((void **)&oq, keep * 4, idx->st));
            GTS_HIP(hipMallocAsync((void **)&on, keep * 4, idx->st));
            GTS_HIP(hipMallocAsync((void **)&om, keep * 4, idx->st));
            GTS_HIP(hipMemcpyAsync(oq, idx->rec_q.p, keep * 4, hipMemcpyDeviceToDevice, idx->st));
            GTS_HIP(hipMemcpyAsync(on, idx->rec_n.p, keep * 4, hipMemcpyDeviceToDevice, idx->st));
            GTS_HIP(hipMemcpyAsync(om, idx->rec_m.p, keep * 4, hipMemcpyDeviceToDevice, idx->st));
            GTS_HIP(hipStreamSynchronize(idx->st));
        }
        idx->rec_q.ensure(need);
        idx->rec_n.ensure(need);
        idx->rec_m.ensure(need);
        if (keep > 0) {
            GTS_HIP(hipMemcpyAsync(idx->rec_q.p, oq, keep * 4, hipMemcpyDeviceToDevice, idx->st));
            GTS_HIP(hipMemcpyAsync(idx->rec_n.p, on, keep * 4, hipMemcpyDeviceToDevice, idx->st));
            GTS_HIP(hipMemcpyAsync(idx->rec_m.p, om, keep * 4, hipMemcpyDeviceToDevice, idx->st));
            GTS_HIP(hipFreeAsync(oq, idx->st));
            GTS_HIP(hipFreeAsync(on, idx->st));
            GTS_HIP(hipFreeAsync(om, idx->st));
        }
    }
}

static int64_t read_scalar(int64_t *d, hipStream_t st)
{
    int64_t v = 0;
    GTS_HIP(hipMemcpyAsync(&v, d, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    GTS_HIP(hipStreamSynchronize(st));
    return v;
}

// Stage 1: frontier expansion; leaves (query, parent, mask) records.
void Search::expand(bool knn)
{
    gts_index *I = idx;
    hipStream_t st = I->st;
    DevTree t = I->dev();
    const bool root_leaf = I->height == 1;
    int64_t *scal = I->scal.ensure(8);
    I->fr_q[0].ensure(nq);
    I->fr_n[0].ensure(nq);
    I->qoff.ensure(nq + 1);
    nrec = 0;
    if (root_leaf) ensure_records(nq);
    k_init_frontier<<<grid_for(nq, 256), 256, 0, st>>>(I->fr_q[0].p, I->fr_n[0].p, I->rec_q.p,
                                                       I->rec_n.p, I->rec_m.p, (int)nq,
                                                       root_leaf ? 1 : 0, I->qoff.p);
    GTS_HIP(hipGetLastError());
    rec_levels = 0;
    rec_pmin = INT32_MAX;
    rec_pmax = -1;
    grouped = v2_ready = false;
    if (root_leaf) {
        nrec = nq;
        rec_levels = 1;
        rec_pmin = rec_pmax = -1;
        return;
    }
    int cur = 0;
    int64_t nf = nq;
    for (int l = 1; l < I->height && nf > 0; l++) {
        int32_t *fq = I->fr_q[cur].p, *fn = I->fr_n[cur].p;
        float *dq = I->fr_dq.ensure(nf);
        stats.frontier_pairs += nf * I->T;
        if (dense_pivot_dist(I, qpos, nq, l - 1, fq, fn, nf, dq)) {
            // dense frontier: one tiled table of all (query, level pivot) distances
        } else if (I->metric == GTS_METRIC_EDIT) {
            k_pivot_dist_edit<<<grid_for(nf, 256, 65536), 256, 0, st>>>(t, fq, fn, nf, qpos, dq);
        } else {
            const int dim = I->dim;
            GTS_METRIC_SWITCH(I->metric, {
                if (dim <= 8)
                    k_pivot_dist<M_, 1><<<grid_for(nf, 256, 65536), 256, 0, st>>>(t, fq, fn, nf, qpos, dq);
                else if (dim <= 128)
                    k_pivot_dist<M_, 16><<<grid_for(nf * 16, 256, 65536), 256, 0, st>>>(t, fq, fn, nf, qpos, dq);
                else
                    k_pivot_dist<M_, 64><<<grid_for(nf * 64, 256, 65536), 256, 0, st>>>(t, fq, fn, nf, qpos, dq);
            });
        }
        GTS_HIP(hipGetLastError());
        if (knn && l >= 3) {
            k_knn_bound<<<(unsigned)nq, 256, 0, st>>>(dq, I->qoff.p, kq, I->pnum[l], (int)nq, bnd, I->em);
            GTS_HIP(hipGetLastError());
        }
        int32_t *ci = I->cnt_int.ensure(nf);
        int32_t *cl = I->qcnt.ensure(nf);   // leaf flags
        uint32_t *lm = I->leafmask.ensure(nf);
        k_children_count<<<grid_for(nf, 256, 65536), 256, 0, st>>>(t, fq, fn, dq, nf, bnd, I->em,
                                                                   ci, cl, lm);
        GTS_HIP(hipGetLastError());
        int64_t *oi = I->off_int.ensure(nf), *ol = I->off_leaf.ensure(nf);
        scan_excl_i32_to_i64(ci, oi, nf, scal + 0, st);
        scan_excl_i32_to_i64(cl, ol, nf, scal + 1, st);
        int64_t tot[2];
        GTS_HIP(hipMemcpyAsync(tot, scal, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, st));
        GTS_HIP(hipStreamSynchronize(st));
        const int64_t nn = tot[0], nl = tot[1];
        // the level's working set: this frontier (pairs, distances, counts,
        // offsets, masks: 40 B a pair), the next frontier, the records so far
        // and their grouped copy, with the buffers' growth slack of 1/4
        budget_check(I, (nf * 40 + nn * 8 + (nrec + nl) * 24) * 5 / 4);
        ensure_records(nrec + nl);
        const int nxt = cur ^ 1;
        I->fr_q[nxt].ensure(nn > 0 ? nn : 1);
        I->fr_n[nxt].ensure(nn > 0 ? nn : 1);
        // fr_q[cur] may have been reallocated? no: only nxt buffers grow here
        k_children_write<<<grid_for(nf, 256, 65536), 256, 0, st>>>(
            t, fq, fn, nf, oi, ol, lm, I->fr_q[nxt].p, I->fr_n[nxt].p, nrec, I->rec_q.p, I->rec_n.p,
            I->rec_m.p);
        GTS_HIP(hipGetLastError());
        k_next_qoff<<<grid_for(nq + 1, 256, 65536), 256, 0, st>>>(I->qoff.p, nq, oi, nf, nn);
        GTS_HIP(hipGetLastError());
        if (nl > 0) {
            rec_levels++;
            rec_pmin = std::min<int32_t>(rec_pmin, (int32_t)I->level_start[l - 1]);
            rec_pmax = std::max<int32_t>(rec_pmax, (int32_t)(I->level_start[l] - 1));
        }
        nrec += nl;
        nf = nn;
        cur = nxt;
    }
    stats.leaf_records += nrec;
}

// Stage 2 preparation: records grouped by parent node, tile offsets.
void Search::group()
{
    if (grouped) return;
    grouped = true;
    gts_index *I = idx;
    hipStream_t st = I->st;
    nslots = I->num_nodes + 1;
    int32_t *cnt = I->node_cnt.ensure(nslots + 1);
    GTS_HIP(hipMemsetAsync(cnt, 0, (nslots + 1) * sizeof(int32_t), st));
    int32_t *gq = I->grp_q.ensure(nrec > 0 ? nrec : 1);
    int32_t *gn = I->grp_n.ensure(nrec > 0 ? nrec : 1);
    uint32_t *gm = I->grp_m.ensure(nrec > 0 ? nrec : 1);
    int64_t *noff = I->node_off.ensure(nslots + 1);
    int64_t *toff = I->tile_off.ensure(nslots + 1);
    int64_t *cur = I->tmp64.ensure(nslots + 1);
    int64_t *scal = I->scal.ensure(8);
    k_node_hist<<<grid_for(nrec, 256, 65536), 256, 0, st>>>(I->rec_n.p, nrec, cnt);
    GTS_HIP(hipGetLastError());
    scan_excl_i32_to_i64(cnt, noff, nslots + 1, scal + 3, st);
    GTS_HIP(hipMemcpyAsync(cur, noff, (nslots + 1) * sizeof(int64_t), hipMemcpyDeviceToDevice, st));
    k_node_scatter<<<grid_for(nrec, 256, 65536), 256, 0, st>>>(
        I->rec_q.p, I->rec_n.p, I->rec_m.p, nrec, (unsigned long long *)cur, gq, gn, gm);
    GTS_HIP(hipGetLastError());
    int32_t *tiles = I->cnt_int.ensure(nslots + 1);
    k_tiles_per_node<<<grid_for(nslots + 1, 256, 65536), 256, 0, st>>>(cnt, nslots + 1, tiles);
    GTS_HIP(hipGetLastError());
    scan_excl_i32_to_i64(tiles, toff, nslots + 1, scal + 4, st);
    ntiles = read_scalar(scal + 4, st);
}

// Exhaustive plan of the BF16 MFMA kNN form: every (query, object) pair goes
// through the dense kernel, filtered by the per-query bound alone.
void Search::verify_all(VerifyArgs &va)
{
    nrec = 0;
    rec_levels = 0;
    stats.exhaustive_passes++;
    if (idx->cells_plan) {
        // the events bracket the verification kernel alone (not the cell
        // bounds); the caller reads them after its next synchronisation
        if (idx->timing) {
            cells_launch(idx, *this, va, idx->timing_ev[0], idx->timing_ev[1], false, 0);
            idx->timing_pending = true;
        } else {
            cells_launch(idx, *this, va, nullptr, nullptr, false, 0);
        }
        return;
    }
    EventTimer tm(idx->st, &stats, idx->timing);
    dense_all_launch(idx, *this, va);
}

void Search::verify(int mode, VerifyArgs &va)
{
    gts_index *I = idx;
    hipStream_t st = I->st;
    DevTree t = I->dev();
    VerifyOut o;
    o.counts = (unsigned long long *)va.counts;
    o.cand = va.cand;
    o.ccount = va.ccount;
    o.cap = va.cap;
    o.band_q = va.band_q;
    o.band_p = va.band_p;
    o.band_n = (unsigned long long *)va.band_n;
    o.band_cap = va.band_cap;
    o.res_k = va.res_k;
    o.res_d = va.res_d;
    o.res_n = (unsigned long long *)va.res_n;
    o.res_cap = va.res_cap;
    o.verified = (unsigned long long *)va.verified;
    if (nrec == 0) return;
    if (v2_supported(I, mode)) {
        static thread_local V2Recs recs;
        if (!v2_ready) {
            recs = v2_prepare(I, *this);
            v2_ready = true;
        }
        EventTimer tm(st, &stats, I->timing);
        v2_launch(I, *this, recs, mode, va);
        return;
    }
    group();
    if (I->metric == GTS_METRIC_EDIT) {
        unsigned g = grid_for(nrec * 64, 256, 1 << 16);
        EventTimer tm(st, &stats, I->timing);
        if (mode == MODE_KNN)
            k_verify_edit<MODE_KNN><<<g, 256, 0, st>>>(t, I->grp_q.p, I->grp_n.p, I->grp_m.p, nrec, qpos, bnd, o);
        else if (mode == MODE_IDS)
            k_verify_edit<MODE_IDS><<<g, 256, 0, st>>>(t, I->grp_q.p, I->grp_n.p, I->grp_m.p, nrec, qpos, bnd, o);
        else
            k_verify_edit<MODE_COUNT><<<g, 256, 0, st>>>(t, I->grp_q.p, I->grp_n.p, I->grp_m.p, nrec, qpos, bnd, o);
        GTS_HIP(hipGetLastError());
        return;
    }
    if (ntiles == 0) return;
    if (ntiles > (int64_t)INT32_MAX) fail(GTS_E_CAPACITY, "too many verification tiles");
    if (mode == MODE_KNN && I->verify_form != GTS_VERIFY_FP32) {
        if (I->metric != GTS_METRIC_L2) fail(GTS_E_ARG, "the BF16 form verifies L2 only");
        ensure_bf16_split(I);
        EventTimer tm(st, &stats, I->timing);
        launch_verify_bf16(I, *this, va);
        return;
    }
    // The expansion's error bound is absolute in the norms: rows of few
    // dimensions in dense clusters away from the origin (2-d points: k-th
    // distances ~1e-5 at norms ~0.7) would hold tens of thousands of rows
    // inside the band, so by default the matrix cores take L2 from
    // FP32_MFMA_MIN_DIM dimensions on, where they pay
    const bool use_mfma = I->fp32_mfma > 0 || (I->fp32_mfma < 0 && I->dim >= FP32_MFMA_MIN_DIM);
    if (I->metric == GTS_METRIC_L2 && use_mfma && mode != MODE_IDS) {
        // L2 counts and kNN candidates on the FP32 matrix cores; the top-k
        // re-ranks with the kernel's error model
        ensure_xn2_max(I);
        const BfErr fe = fp32_mfma_err_model(I->dp);
        int32_t *tslot = I->tile_slot.ensure(ntiles);
        k_tile_slot<<<grid_for(nslots, 256, 65536), 256, 0, st>>>(I->tile_off.p, nslots, tslot);
        GTS_HIP(hipGetLastError());
        const unsigned grid = (unsigned)ntiles;
        EventTimer tm(st, &stats, I->timing);
        if (mode == MODE_KNN) {
            k_verify_mfma<MODE_KNN><<<grid, 256, 0, st>>>(t, I->grp_q.p, I->grp_m.p, I->node_off.p, I->tile_off.p,
                                                         tslot, ntiles, qpos, bnd, fe, o);
            mfma_keys = true;
        } else {
            k_verify_mfma<MODE_COUNT><<<grid, 256, 0, st>>>(t, I->grp_q.p, I->grp_m.p, I->node_off.p, I->tile_off.p,
                                                           tslot, ntiles, qpos, bnd, fe, o);
        }
        GTS_HIP(hipGetLastError());
        return;
    }
    EventTimer tm(st, &stats, I->timing);
    GTS_METRIC_SWITCH(I->metric, {
        if (mode == MODE_KNN)
            k_verify<M_, MODE_KNN><<<(unsigned)ntiles, 256, 0, st>>>(t, I->grp_q.p, I->grp_m.p, I->node_off.p,
                                                                    I->tile_off.p, nslots, qpos, bnd, I->em, o);
        else if (mode == MODE_IDS)
            k_verify<M_, MODE_IDS><<<(unsigned)ntiles, 256, 0, st>>>(t, I->grp_q.p, I->grp_m.p, I->node_off.p,
                                                                    I->tile_off.p, nslots, qpos, bnd, I->em, o);
        else
            k_verify<M_, MODE_COUNT><<<(unsigned)ntiles, 256, 0, st>>>(t, I->grp_q.p, I->grp_m.p, I->node_off.p,
                                                                      I->tile_off.p, nslots, qpos, bnd, I->em, o);
    });
    GTS_HIP(hipGetLastError());
}

void Search::band(int mode, VerifyArgs &va, int64_t nb)
{
    if (nb <= 0 || idx->metric == GTS_METRIC_EDIT) return;
    gts_index *I = idx;
    DevTree t = I->dev();
    VerifyOut o{};
    o.counts = (unsigned long long *)va.counts;
    o.res_k = va.res_k;
    o.res_d = va.res_d;
    o.res_n = (unsigned long long *)va.res_n;
    o.res_cap = va.res_cap;
    unsigned g = grid_for(nb * 64, 256, 1 << 16);
    GTS_METRIC_SWITCH(I->metric, {
        if (mode == MODE_IDS)
            k_band_range<M_, MODE_IDS><<<g, 256, 0, I->st>>>(t, va.band_q, va.band_p, nb, qpos, bnd, o);
        else
            k_band_range<M_, MODE_COUNT><<<g, 256, 0, I->st>>>(t, va.band_q, va.band_p, nb, qpos, bnd, o);
    });
    GTS_HIP(hipGetLastError());
}

void Search::topk(VerifyArgs &va, int check_radius, const int32_t *orig, int kstride, int32_t *ids,
                  float *dists, float *kth, int32_t *status, int32_t *failq, unsigned long long *ctr,
                  bool count_candidates)
{
    // candidate totals here only when the verification did not count them:
    // one atomic per query on one address
    unsigned long long *ncand = (ctr && count_candidates) ? ctr + 1 : nullptr;
    gts_index *I = idx;
    DevTree t = I->dev();
    if (va.cap > TOPK_CAP) fail(GTS_E_ARG, "candidate capacity above the top-k limit");
    const int bf = ((I->verify_form != GTS_VERIFY_FP32 || mfma_keys) && I->metric == GTS_METRIC_L2) ? 1 : 0;
    BfErr be = mfma_keys ? fp32_mfma_err_model(I->dp) : bf_err_model(I->dim, I->dk, 3);
    // the cell scan's rows carry FP8 residuals: their largest row error
    if (I->cells_plan) be.ex = I->cells.eps_max;
    // k <= 64: the wave-level kernel, then the block-level one for the queries
    // it deferred; larger k or a small batch (reruns: few queries, long
    // candidate lists that one wave would walk 64 at a time): the block-level
    // kernel for every query
    const bool wave = kstride <= WAVE_K && nq >= 256;
    const unsigned wgrid = (unsigned)((nq + 3) / 4);
    if (I->metric == GTS_METRIC_EDIT) {
        if (wave)
            k_topk_wave<GTS_METRIC_EDIT><<<wgrid, 256, 0, I->st>>>(t, I->inv, va.cand, va.ccount, va.cap, qpos, kq,
                                                                   bnd, check_radius, I->em, 0, be, I->xn2,
                                                                   I->xn2_max, orig, kstride, ids, dists, kth,
                                                                   status, nq, failq, ctr, ncand);
        k_topk<GTS_METRIC_EDIT><<<(unsigned)nq, 256, 0, I->st>>>(t, I->inv, va.cand, va.ccount, va.cap, qpos, kq,
                                                                 bnd, check_radius, I->em, 0, be, I->xn2,
                                                                 I->xn2_max, orig, kstride, ids, dists, kth,
                                                                 status, wave ? 1 : 0, failq, ctr, ncand);
    } else {
        GTS_METRIC_SWITCH(I->metric, {
            if (wave)
                k_topk_wave<M_><<<wgrid, 256, 0, I->st>>>(t, I->inv, va.cand, va.ccount, va.cap, qpos, kq, bnd,
                                                          check_radius, I->em, bf, be, I->xn2, I->xn2_max, orig,
                                                          kstride, ids, dists, kth, status, nq, failq, ctr, ncand);
            k_topk<M_><<<(unsigned)nq, 256, 0, I->st>>>(t, I->inv, va.cand, va.ccount, va.cap, qpos, kq, bnd,
                                                        check_radius, I->em, bf, be, I->xn2, I->xn2_max, orig,
                                                        kstride, ids, dists, kth, status, wave ? 1 : 0, failq, ctr, ncand);
        });
    }
    GTS_HIP(hipGetLastError());
}

int topk_cap_limit() { return TOPK_CAP; }

namespace {

__global__ void k_stride_offsets(int64_t *off, int64_t n, int64_t stride)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n) off[i] = i * stride;
}

constexpr int64_t RP_CH = 16384;   // queries per pivot-radius table (<= 16384 x 4096 x 4 B = 256 MB)

}  // namespace

std::vector<int32_t> level_pivots(gts_index *I, int maxp, int *level, int minp)
{
    std::vector<uint8_t> fl(I->num_nodes);
    std::vector<int32_t> pv(I->num_nodes);
    GTS_HIP(hipMemcpyAsync(fl.data(), I->nflag, I->num_nodes, hipMemcpyDeviceToHost, I->st));
    GTS_HIP(hipMemcpyAsync(pv.data(), I->npiv, I->num_nodes * 4, hipMemcpyDeviceToHost, I->st));
    GTS_HIP(hipStreamSynchronize(I->st));
    std::vector<int32_t> best;
    *level = -1;
    for (int l = 0; l + 1 < I->height; l++) {
        std::vector<int32_t> cur;
        for (int64_t nd = I->level_start[l]; nd < I->level_start[l + 1]; nd++)
            if ((fl[nd] & 3) == 1) cur.push_back(pv[nd]);   // existing split node
        if ((int)cur.size() > maxp) break;
        if (cur.size() >= best.size()) {
            best = cur;
            *level = l;
        }
        if ((int)best.size() >= minp) break;   // the shallowest level with enough pivots
    }
    return best;
}

void pivot_radius_prepare(gts_index *I)
{
    if (I->rp_n >= 0) return;
    if (I->metric == GTS_METRIC_EDIT) fail(GTS_E_ARG, "the pivot radius needs a vector metric");
    int level = -1;
    const std::vector<int32_t> v = level_pivots(I, RADIUS_MAXP, &level);
    int32_t *d = I->rp_piv.ensure(std::max<int64_t>((int64_t)v.size(), 1));
    if (!v.empty()) GTS_HIP(hipMemcpyAsync(d, v.data(), v.size() * 4, hipMemcpyHostToDevice, I->st));
    I->rp_n = (int64_t)v.size();
    I->rp_level = level;
}

void pivot_radius_buffers(gts_index *I, int64_t nq)
{
    pivot_radius_prepare(I);
    const int64_t ch = std::min(nq, RP_CH);
    I->ftab.ensure(std::max<int64_t>(ch * I->rp_n, 1));
    I->rp_off.ensure(ch + 1);
}

void pivot_radius(gts_index *I, const int32_t *qpos, const int32_t *kq, int64_t nq, float *bnd)
{
    pivot_radius_buffers(I, nq);
    hipStream_t st = I->st;
    fill_f32(bnd, INFINITY, nq, st);
    const int64_t P = I->rp_n;
    if (P == 0 || nq == 0) return;
    float *D = I->ftab.ensure(std::min(nq, RP_CH) * P);
    int64_t *off = I->rp_off.ensure(std::min(nq, RP_CH) + 1);
    for (int64_t q0 = 0; q0 < nq; q0 += RP_CH) {
        const int64_t m = std::min(RP_CH, nq - q0);
        pivot_table(I, qpos + q0, m, I->rp_piv.p, P, D);
        k_stride_offsets<<<grid_for(m + 1, 256), 256, 0, st>>>(off, m, P);
        k_knn_bound<<<(unsigned)m, 256, 0, st>>>(D, off, kq + q0, P, (int)m, bnd + q0, I->em);
        GTS_HIP(hipGetLastError());
    }
}

}  // namespace gts
