// Query-block-stationary verification (FP32 VALU and BF16 hi+lo MFMA).
//
// Replaces the per-pair leaf stage of the reference (dataProcessRnn /
// dataProcessKnn, search_v2.cuh:319-544) for dense workloads.  A block keeps
// 64 consecutive queries resident in LDS and walks their leaf records
// (query, parent, leaf-mask), which are sorted by parent within each query,
// as a 64-way merge over one range of parent IDs: at each step the smallest
// pending parent P is taken, every query whose list holds P contributes its
// leaf mask, and the objects of the masked leaf children of P (contiguous
// positions) are verified against all 64 queries at once.  Query rows are
// therefore read once per block instead of once per parent, and blocks that
// share a parent range run together so object rows are served from L2/MALL.
#include <climits>

#include "gts_dist.h"
#include "gts_index.h"
#include "gts_mfma.h"
#include "gts_wave.h"
#include "gts_search.h"
#include "gts_verify2.h"

namespace gts {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int QB = 64;    // queries per block
constexpr int RT = 128;   // objects per tile
constexpr int DKF = 32;   // dims per LDS stage of object rows (FP32 kernel)

// shared merge state of one block
struct Merge {
    int64_t cur[QB], end[QB];
    int32_t q[QB], qp[QB];
    int32_t pend[QB];
    uint32_t mask[QB];
    float b[QB];
    float qn[QB];
    int32_t P;
    uint32_t umask;
    int32_t cnt[QB];
    int32_t base[QB];
    uint32_t tile_n;
    unsigned long long tile_base;
};

__device__ __forceinline__ int64_t lower_bound_i32(const int32_t *a, int64_t lo, int64_t hi, int32_t v)
{
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// block setup: queries of block qb, cursors at the first record with parent >= Pa
__device__ void merge_init(Merge &S, const int32_t *rec_n, const int64_t *roff, int64_t nq,
                           int64_t qb, int32_t Pa, const int32_t *qpos, const float *bnd,
                           const float *xn2)
{
    const int tid = threadIdx.x;
    if (tid < QB) {
        int64_t q = qb * QB + tid;
        if (q < nq) {
            int64_t s = roff[q], e = roff[q + 1];
            S.cur[tid] = lower_bound_i32(rec_n, s, e, Pa);
            S.end[tid] = e;
            S.q[tid] = (int32_t)q;
            S.qp[tid] = qpos[q];
            S.b[tid] = bnd[q];
            S.qn[tid] = xn2 ? xn2[qpos[q]] : 0.0f;
        } else {
            S.cur[tid] = 0;
            S.end[tid] = 0;
            S.q[tid] = -1;
            S.qp[tid] = 0;
            S.b[tid] = -1.0f;
            S.qn[tid] = 0.0f;
        }
    }
}

// next parent of the merge (all threads); false when every list is exhausted
__device__ bool merge_next(Merge &S, const int32_t *rec_n, const uint32_t *rec_m, int32_t Pb)
{
    const int tid = threadIdx.x;
    __syncthreads();  // every wave is done with the previous parent
    if (tid < QB) {
        int32_t p = INT_MAX;
        if (S.cur[tid] < S.end[tid]) {
            p = rec_n[S.cur[tid]];
            if (p >= Pb) p = INT_MAX;
        }
        S.pend[tid] = p;
        // wave 0 holds all 64 queries: reduce the minimum with shuffles
        int32_t m = p;
        for (int o = 32; o > 0; o >>= 1) m = min(m, __shfl_xor(m, o, 64));
        if (tid == 0) {
            S.P = m;
            S.umask = 0;
        }
    }
    __syncthreads();
    const int32_t P = S.P;
    if (P == INT_MAX) return false;
    if (tid < QB) {
        uint32_t mk = 0;
        if (S.pend[tid] == P) {
            mk = rec_m[S.cur[tid]];
            S.cur[tid]++;
        }
        S.mask[tid] = mk;
        if (mk) atomicOr(&S.umask, mk);
    }
    __syncthreads();
    return true;
}

__device__ __forceinline__ void run_range(const DevTree &t, int32_t P, int c, int cend,
                                          int32_t &clid, int32_t &run)
{
    if (P < 0) {
        clid = 0;
        run = (int32_t)t.n;
    } else {
        int64_t c0 = (int64_t)P * t.T + 1;
        clid = t.nlid[c0 + c];
        run = t.nlid[c0 + cend] + t.nsize[c0 + cend] - clid;
    }
}

__device__ __forceinline__ int child_of(const DevTree &t, int32_t P, int32_t plid, int32_t pavg, int32_t pos)
{
    if (P < 0) return 0;
    return pavg > 0 ? min((pos - plid) / pavg, t.T - 1) : t.T - 1;
}

// ---------------------------------------------------------------- BF16 MFMA
__global__ void __launch_bounds__(256) k_verify2_bf16(DevTree t, const uint16_t *__restrict__ Xh,
                                                      const uint16_t *__restrict__ Xl,
                                                      const float *__restrict__ xn2, int dk,
                                                      const int32_t *rec_n, const uint32_t *rec_m,
                                                      const int64_t *roff, int64_t nq, int64_t nqb,
                                                      int32_t pmin, int32_t pspan, int nranges,
                                                      const int32_t *qpos, const float *bnd,
                                                      BfErr be, uint64_t *cand, int32_t *ccount,
                                                      int cap, unsigned long long *verified)
{
    extern __shared__ __attribute__((aligned(16))) uint16_t smem[];
    __shared__ Merge S;
    const int pitch = dk + 8;
    uint16_t *Qh = smem, *Ql = smem + QB * pitch;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int64_t qb = blockIdx.x % nqb;
    const int rg = (int)(blockIdx.x / nqb);
    const int32_t Pa = pmin + (int32_t)((int64_t)pspan * rg / nranges);
    const int32_t Pb = pmin + (int32_t)((int64_t)pspan * (rg + 1) / nranges);
    merge_init(S, rec_n, roff, nq, qb, Pa, qpos, bnd, xn2);
    __syncthreads();
    // resident query rows (hi and lo parts)
    const int cpr = dk / 8;  // 16-byte chunks per row
    for (int c = tid; c < QB * cpr; c += blockDim.x) {
        const int i = c / cpr, k8 = (c % cpr) * 8;
        uint4 h = make_uint4(0, 0, 0, 0), l = h;
        if (S.q[i] >= 0) {
            const int64_t o = frag_off(S.qp[i], k8, cpr / 4);
            h = *(const uint4 *)(Xh + o);
            l = *(const uint4 *)(Xl + o);
        }
        *(uint4 *)(Qh + i * pitch + k8) = h;
        *(uint4 *)(Ql + i * pitch + k8) = l;
    }
    __syncthreads();
    unsigned long long nver = 0;
    while (merge_next(S, rec_n, rec_m, Pb)) {
        const int32_t P = S.P;
        const uint32_t um = S.umask;
        const int nchild = P < 0 ? 1 : t.T;
        const int32_t plid = P < 0 ? 0 : t.nlid[P];
        const int32_t pavg = P < 0 ? 0 : t.nsize[P] / t.T;
        for (int c = 0; c < nchild; c++) {
            if (!((um >> c) & 1u)) continue;
            int cend = c;
            while (cend + 1 < nchild && ((um >> (cend + 1)) & 1u)) cend++;
            int32_t clid, run;
            run_range(t, P, c, cend, clid, run);
            c = cend;
            for (int r0 = 0; r0 < run; r0 += RT) {
                f32x4 acc[4][2];
#pragma unroll
                for (int a = 0; a < 4; a++)
#pragma unroll
                    for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
                // this lane's object rows for the two 16-row MFMA blocks of its wave
                int64_t rowoff[2];
                bool rv[2];
#pragma unroll
                for (int b = 0; b < 2; b++) {
                    const int j = w * 32 + 16 * b + (lane & 15);
                    rv[b] = r0 + j < run;
                    rowoff[b] = clid + r0 + (rv[b] ? j : 0);
                }
                const int kl = 8 * (lane >> 4);
                for (int k0 = 0; k0 < dk; k0 += 32) {
                    bf16x8 bh[2], bl[2], ah[4], al[4];
#pragma unroll
                    for (int b = 0; b < 2; b++) {
                        uint4 h = make_uint4(0, 0, 0, 0), l = h;
                        if (rv[b]) {
                            const int64_t o = frag_off(rowoff[b], k0 + 8 * (lane >> 4), dk / 32);
                            h = *(const uint4 *)(Xh + o);
                            l = *(const uint4 *)(Xl + o);
                        }
                        bh[b] = __builtin_bit_cast(bf16x8, h);
                        bl[b] = __builtin_bit_cast(bf16x8, l);
                    }
#pragma unroll
                    for (int a = 0; a < 4; a++) {
                        const int qi = 16 * a + (lane & 15);
                        ah[a] = *(const bf16x8 *)(Qh + qi * pitch + k0 + kl);
                        al[a] = *(const bf16x8 *)(Ql + qi * pitch + k0 + kl);
                    }
#pragma unroll
                    for (int a = 0; a < 4; a++)
#pragma unroll
                        for (int b = 0; b < 2; b++) {
                            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
                            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
                            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
                        }
                }
                // epilogue: lane holds q = 16a + 4(lane/16) + r, object row w*32 + 16b + lane%16
                if (tid < QB) S.cnt[tid] = 0;
                __syncthreads();
                int slot[4][2][4];
                uint32_t hit = 0;
#pragma unroll
                for (int a = 0; a < 4; a++)
#pragma unroll
                    for (int b = 0; b < 2; b++)
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                            const int bit = a * 8 + b * 4 + r;
                            slot[a][b][r] = 0;
                            const int i = 16 * a + 4 * (lane >> 4) + r;
                            const int j = w * 32 + 16 * b + (lane & 15);
                            if (!rv[b]) continue;
                            const int32_t pj = clid + r0 + j;
                            if (!((S.mask[i] >> child_of(t, P, plid, pavg, pj)) & 1u)) continue;
                            if (hidden(t, pj, S.q[i])) continue;
                            nver++;
                            const float qn = S.qn[i], xn = xn2[pj];
                            const float d = sqrtf(fmaxf(qn + xn - 2.0f * acc[a][b][r], 0.0f));
                            const float E2 = be.c1 * (qn + xn) + 2.0f * be.c2 * sqrtf(qn * xn);
                            const float e = fminf(sqrtf(E2), d > 0.0f ? E2 / d : INFINITY);
                            if (S.qp[i] == pj || d - e <= S.b[i]) {
                                hit |= 1u << bit;
                                slot[a][b][r] = atomicAdd(&S.cnt[i], 1);
                            }
                        }
                __syncthreads();
                if (tid < QB && S.cnt[tid] > 0) S.base[tid] = atomicAdd(&ccount[S.q[tid]], S.cnt[tid]);
                __syncthreads();
                if (hit) {
#pragma unroll
                    for (int a = 0; a < 4; a++)
#pragma unroll
                        for (int b = 0; b < 2; b++)
#pragma unroll
                            for (int r = 0; r < 4; r++) {
                                if (!(hit & (1u << (a * 8 + b * 4 + r)))) continue;
                                const int i = 16 * a + 4 * (lane >> 4) + r;
                                const int32_t pj = clid + r0 + w * 32 + 16 * b + (lane & 15);
                                const float qn = S.qn[i], xn = xn2[pj];
                                float d = sqrtf(fmaxf(qn + xn - 2.0f * acc[a][b][r], 0.0f));
                                if (S.qp[i] == pj) d = 0.0f;
                                const int s = S.base[i] + slot[a][b][r];
                                if (s < cap)
                                    cand[(int64_t)S.q[i] * cap + s] =
                                        ((uint64_t)__float_as_uint(d) << 32) | (uint32_t)t.perm[pj];
                            }
                }
                __syncthreads();
            }
        }
    }
    block_add_u64(nver, verified);
}

// ------------------------------------------------------ dense BF16 MFMA
enum { V2_COUNT = 0, V2_IDS = 1, V2_KNN = 2 };

constexpr int DW = 12;    // waves per block of the dense kernel (3 per SIMD)
constexpr int DR = 32;    // object rows per wave chunk

// LDS image of every 64-query block of the batch for the dense kernel: BF16
// hi rows then lo rows, pitch dk + 8, split exactly as the object copies
// (k_split_bf16), zero rows for missing queries.  Blocks of the dense kernel
// copy their image with contiguous loads.  A thread converts 8 consecutive
// values (two 16-byte loads of the FP32 row, zero-padded up to dp, a multiple
// of 4; one 16-byte store each of hi and lo; pitch is a multiple of 8).
__global__ void k_stage_queries(const float *__restrict__ X, int dim, int dp, int dk, const int32_t *qpos,
                                int64_t nq, int64_t nqb, uint16_t *img)
{
    const int pitch = dk + 8, pv = pitch / 8;
    const int64_t per = (int64_t)2 * QB * pitch;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nqb * QB * pv;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t qb = i / ((int64_t)QB * pv);
        const int r = (int)(i / pv % QB), k0 = (int)(i % pv) * 8;
        const int64_t q = qb * QB + r;
        float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f), b = a;
        if (q < nq) {
            const float *src = X + (int64_t)qpos[q] * dp + k0;
            if (k0 < dp) a = *(const float4 *)src;
            if (k0 + 4 < dp) b = *(const float4 *)(src + 4);
        }
        const float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        uint32_t h[4], l[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const float v0 = k0 + 2 * j < dim ? v[2 * j] : 0.0f, v1 = k0 + 2 * j + 1 < dim ? v[2 * j + 1] : 0.0f;
            const uint16_t h0 = bf16_rn(v0), h1 = bf16_rn(v1);
            const uint16_t l0 = bf16_rn(v0 - __uint_as_float((uint32_t)h0 << 16));
            const uint16_t l1 = bf16_rn(v1 - __uint_as_float((uint32_t)h1 << 16));
            h[j] = (uint32_t)h0 | ((uint32_t)h1 << 16);
            l[j] = (uint32_t)l0 | ((uint32_t)l1 << 16);
        }
        uint16_t *o = img + qb * per + (int64_t)r * pitch + k0;
        *(uint4 *)o = make_uint4(h[0], h[1], h[2], h[3]);
        *(uint4 *)(o + (int64_t)QB * pitch) = make_uint4(l[0], l[1], l[2], l[3]);
    }
}

// Dense regime: nearly every (query, parent) pair of one level survived, so a
// block streams a contiguous range of object rows past 64 resident queries.
// Each wave owns every DW-th 32-row chunk of the range and runs its K steps
// with no block barrier.  The B fragments (object rows, hi and lo) of all the
// wave's chunks form one stream of K steps, loaded two steps ahead of the
// MFMAs that consume them, across chunk boundaries.  Per-row epilogue data
// (norm terms, parent node) of the block's range is staged in LDS once, so the
// epilogue issues no global load outside the candidate path.  The leaf mask of
// a (query, row) pair is the record mask of (query, parent of row) in a dense
// table; without a table (maskT null) every row is a pair of every query
// (exhaustive plan).
template <bool PERSIST>
__global__ void __launch_bounds__(DW * 64) k_verify_dense_bf16(
    DevTree t, const uint16_t *__restrict__ Xh, const uint16_t *__restrict__ Xl,
    const float *__restrict__ xn2, int dk, const uint32_t *__restrict__ maskT, int64_t nP,
    int32_t pmin, const int32_t *__restrict__ par_of, int64_t nq, int64_t nqb, int nranges,
    const uint16_t *__restrict__ qimg, const int32_t *qpos, const float *bnd, BfErr be, uint64_t *cand,
    int32_t *ccount, int cap, unsigned long long *verified)
{
    extern __shared__ __attribute__((aligned(16))) uint16_t smem[];
    __shared__ int32_t s_q[QB], s_qp[QB];
    __shared__ float s_qn[QB];
    __shared__ __attribute__((aligned(16))) float s_tq[QB], s_cq[QB];
    const int pitch = dk + 8;
    uint16_t *Qh = smem, *Ql = smem + QB * pitch;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave index, uniform in the wave
    const int64_t n = t.n;
    unsigned long long nver = 0;
    // One work item = (query block qb, row range rg).  new_qb: the block's LDS
    // does not hold qb's query image yet.
    auto process = [&](const int64_t qb, const int rg, const bool new_qb) {

    const int64_t nchunk = (n + DR - 1) / DR;
    const int64_t ca = nchunk * rg / nranges, cb = nchunk * (rg + 1) / nranges;
    const int64_t rbeg = ca * DR;
    const int nrow = (int)((cb - ca) * DR);   // rows of the range, padding included
    float *s_ux = (float *)(smem + 2 * QB * pitch), *s_xr = s_ux + nrow;
    int32_t *s_par = (int32_t *)(s_xr + nrow);
    const float c1 = be.c1 + 8.0f * 5.9604645e-08f;
    // Candidate filter per pair: keep iff the exact d^2 may be <= b^2, i.e.
    //   d2 - E2 <= b^2  with  d2 = |q|^2 + |x|^2 - 2 q.x,  E2 = c1 (|q|^2 + |x|^2) + 2 c2 |q||x|
    //   <=>  q.x >= T_q + U_x - c2 |q| |x|,
    //   T_q = ((1 - c1) |q|^2 - b^2) / 2,  U_x = (1 - c1) |x|^2 / 2.
    // T_q and c2 |q| of the block's queries sit in LDS (s_tq, s_cq).
    if (new_qb && tid < QB) {
        const int64_t q = qb * QB + tid;
        const bool ok = q < nq;
        const float qn = ok ? xn2[qpos[q]] : 0.0f;
        const float b = ok ? bnd[q] : -1.0f;
        s_q[tid] = ok ? (int32_t)q : -1;
        s_qp[tid] = ok ? qpos[q] : 0;
        s_qn[tid] = qn;
        s_tq[tid] = ok ? 0.5f * ((1.0f - c1) * qn - b * b * 1.000001f) : INFINITY;
        s_cq[tid] = be.c2 * sqrtf(qn);
    }
    // Staging loads are issued four per thread before their LDS stores, so a
    // block waits about two memory round trips instead of one per iteration.
    for (int i0 = tid; i0 < nrow; i0 += 4 * blockDim.x) {
        float xv[4];
        int32_t pv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int i = i0 + u * blockDim.x;
            const int64_t row = rbeg + i;
            const bool ok = i < nrow && row < n;
            xv[u] = ok ? xn2[row] : 0.0f;
            pv[u] = ok ? (maskT ? par_of[row] : 0) : -1;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int i = i0 + u * blockDim.x;
            if (i < nrow) {
                s_ux[i] = 0.5f * (1.0f - c1) * xv[u];
                s_xr[i] = sqrtf(xv[u]);
                s_par[i] = pv[u];
            }
        }
    }
    if (new_qb) {
        // the query block's hi/lo image: 2 * QB * pitch BF16, a multiple of 8
        const uint4 *src = (const uint4 *)(qimg + qb * 2 * QB * pitch);
        uint4 *dst = (uint4 *)smem;
        const int nimg = QB * pitch / 4;
        for (int c0 = tid; c0 < nimg; c0 += 4 * blockDim.x) {
            uint4 v[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int c = c0 + u * blockDim.x;
                v[u] = c < nimg ? src[c] : make_uint4(0, 0, 0, 0);
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int c = c0 + u * blockDim.x;
                if (c < nimg) dst[c] = v[u];
            }
        }
    }
    __syncthreads();
    const int kl = 8 * (lane >> 4);
    const int nk = dk / 32;
    int nvq = 0;
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int r = 0; r < 4; r++) nvq += s_q[16 * a + 4 * (lane >> 4) + r] >= 0;
    // The wave's chunks are c_first + j DW, j < nmy; its K-step stream is
    // s = j nk + k.  In the fragment-major layout the lane's fragment of step
    // (j, k) for the chunk's first 16-row block sits at element
    //   base + (j (2 DW nk) + k) * 512,   base = (c_first * 2 nk * 64 + lane) * 8,
    // and the second block's 16 rows follow nk * 512 elements later.
    const int64_t c_first = ca + w;
    const int64_t nmy = c_first < cb ? (cb - 1 - c_first) / DW + 1 : 0;
    const int64_t nsteps = nmy * nk;
    const int64_t jstride = (int64_t)2 * DW * nk;             // fragment steps between the wave's chunks
    const int64_t flast = (nmy - 1) * jstride + (nk - 1);     // last fragment step of the stream
    const uint16_t *hb = Xh + (c_first * 2 * nk * 64 + lane) * 8;
    const uint16_t *lb = Xl + (c_first * 2 * nk * 64 + lane) * 8;
    const int64_t blk2 = (int64_t)nk * 512;
    f32x4 acc[4][2];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    // Loader state of the three buffers X, Y, Z (steps s = 0, 1, 2 mod 3):
    // fragment step f and K step k of the next load.  Loads past the end of the
    // stream re-read its last step, so every load is unconditional and the
    // wait counts stay exact; sched_barrier keeps the scheduler from sinking
    // loads to their uses.
    int64_t fx = 0, fy = 1, fz = 2;
    int kx = 0, ky = 1, kz = 2;
    auto norm = [&](int64_t &f, int &k) {
        while (k >= nk) {
            k -= nk;
            f += jstride - nk;
        }
    };
    norm(fy, ky);
    norm(fz, kz);
    auto ld = [&](int64_t &f, int &k, uint4 &a0, uint4 &a1, uint4 &b0, uint4 &b1) {
        const int64_t o = min(f, flast) * 512;
        a0 = *(const uint4 *)(hb + o);
        a1 = *(const uint4 *)(hb + o + blk2);
        b0 = *(const uint4 *)(lb + o);
        b1 = *(const uint4 *)(lb + o + blk2);
        f += 3;
        k += 3;
        norm(f, k);
    };
    // consumer state: chunk ordinal j and K step k of the next step
    int64_t cj = 0;
    int ck = 0;
    // A fragments of the 4 query blocks (hi and lo) are read from LDS at the
    // start of each step; the other waves of the SIMD cover their latency.
    auto step = [&](const uint4 &a0, const uint4 &a1, const uint4 &b0, const uint4 &b1) {
        bf16x8 ah[4], al[4];
#pragma unroll
        for (int a = 0; a < 4; a++) {
            const int qi = 16 * a + (lane & 15);
            ah[a] = *(const bf16x8 *)(Qh + qi * pitch + 32 * ck + kl);
            al[a] = *(const bf16x8 *)(Ql + qi * pitch + 32 * ck + kl);
        }
        const bf16x8 vh[2] = {__builtin_bit_cast(bf16x8, a0), __builtin_bit_cast(bf16x8, a1)};
        const bf16x8 vl[2] = {__builtin_bit_cast(bf16x8, b0), __builtin_bit_cast(bf16x8, b1)};
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 2; b++)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], vh[b], acc[a][b], 0, 0, 0);
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 2; b++)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], vl[b], acc[a][b], 0, 0, 0);
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 2; b++)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], vh[b], acc[a][b], 0, 0, 0);
        // all eight LDS reads ahead of the 24 MFMAs
        __builtin_amdgcn_sched_group_barrier(0x100, 8, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 24, 0);
        if (ck == nk - 1) {
            float tq[4][4], cq[4][4];
#pragma unroll
            for (int a = 0; a < 4; a++) {
                const float4 t4 = *(const float4 *)&s_tq[16 * a + 4 * (lane >> 4)];
                const float4 c4 = *(const float4 *)&s_cq[16 * a + 4 * (lane >> 4)];
                tq[a][0] = t4.x; tq[a][1] = t4.y; tq[a][2] = t4.z; tq[a][3] = t4.w;
                cq[a][0] = c4.x; cq[a][1] = c4.y; cq[a][2] = c4.z; cq[a][3] = c4.w;
            }
            // epilogue of chunk cj: lane holds query 16a + 4(lane/16) + r and
            // row r0 + 16b + lane%16; three VALU operations per pair, the rest
            // only for candidates
            const int64_t r0 = (c_first + cj * DW) * DR;
            float ux[2], xr[2];
            int32_t P[2];
            bool hit = false;
#pragma unroll
            for (int b = 0; b < 2; b++) {
                const int li = (int)(r0 - rbeg) + 16 * b + (lane & 15);
                ux[b] = s_ux[li];
                xr[b] = s_xr[li];
                P[b] = s_par[li];
                if (P[b] >= 0) nver += nvq;
                bool h = false;
#pragma unroll
                for (int a = 0; a < 4; a++)
#pragma unroll
                    for (int r = 0; r < 4; r++) h |= acc[a][b][r] >= fmaf(-cq[a][r], xr[b], tq[a][r] + ux[b]);
                hit |= h && P[b] >= 0;
            }
            // candidate path (rare): only waves with a candidate re-test their pairs
#pragma unroll
            for (int b = 0; b < 2 && __builtin_amdgcn_ballot_w64(hit); b++) {
                const int64_t row = r0 + 16 * b + (lane & 15);
#pragma unroll
                for (int a = 0; a < 4; a++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const float rhs = fmaf(-cq[a][r], xr[b], tq[a][r] + ux[b]);
                        if (!(acc[a][b][r] >= rhs) || P[b] < 0) continue;
                        // leaf mask, tombstone, append
                        const int i = 16 * a + 4 * (lane >> 4) + r;
                        const int32_t q = s_q[i];
                        if (maskT) {
                            const int32_t plid = t.nlid[P[b]];
                            const int32_t pavg = t.nsize[P[b]] / t.T;
                            const int cjx = pavg > 0 ? min((int32_t)(row - plid) / pavg, t.T - 1) : t.T - 1;
                            if (!((maskT[(int64_t)q * nP + (P[b] - pmin)] >> cjx) & 1u)) continue;
                        }
                        if (hidden(t, row, q)) continue;
                        const bool self = s_qp[i] == (int32_t)row;
                        const float d = self ? 0.0f : sqrtf(fmaxf(s_qn[i] + xn2[row] - 2.0f * acc[a][b][r], 0.0f));
                        const int sl = atomicAdd(&ccount[q], 1);
                        if (sl < cap)
                            cand[(int64_t)q * cap + sl] = ((uint64_t)__float_as_uint(d) << 32) | (uint32_t)t.perm[row];
                    }
            }
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
            ck = 0;
            cj++;
        } else {
            ck++;
        }
    };
    if (nsteps > 0) {
        uint4 xh0, xh1, xl0, xl1, yh0, yh1, yl0, yl1, zh0, zh1, zl0, zl1;
        ld(fx, kx, xh0, xh1, xl0, xl1);
        ld(fy, ky, yh0, yh1, yl0, yl1);
        ld(fz, kz, zh0, zh1, zl0, zl1);
        __builtin_amdgcn_sched_barrier(0);
        for (int64_t s0 = 0; s0 < nsteps; s0 += 3) {
            step(xh0, xh1, xl0, xl1);
            __builtin_amdgcn_sched_barrier(0);
            ld(fx, kx, xh0, xh1, xl0, xl1);
            __builtin_amdgcn_sched_barrier(0);
            if (s0 + 1 < nsteps) step(yh0, yh1, yl0, yl1);
            __builtin_amdgcn_sched_barrier(0);
            ld(fy, ky, yh0, yh1, yl0, yl1);
            __builtin_amdgcn_sched_barrier(0);
            if (s0 + 2 < nsteps) step(zh0, zh1, zl0, zl1);
            __builtin_amdgcn_sched_barrier(0);
            ld(fz, kz, zh0, zh1, zl0, zl1);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    };
    if (!PERSIST) {
        // XCD-aware mapping: blocks are dispatched round-robin over the 8 XCDs,
        // so hardware block b runs on XCD b % 8.  The grid is padded to a
        // multiple of 8 and logical block L = (b % 8) * (grid / 8) + b / 8 (a
        // bijection), so consecutive logical blocks -- the query blocks of one
        // row range -- share an XCD and reuse that range's rows from its L2.
        const int64_t per = (int64_t)gridDim.x / 8;
        const int64_t L = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
        if (L >= nqb * (int64_t)nranges) return;
        process(L % nqb, (int)(L / nqb), true);
    } else {
        // Persistent blocks, one per CU, blocks b % 8 on XCD b % 8 (the same
        // dispatch order).  XCD x owns row ranges [x R / 8, (x + 1) R / 8);
        // its blocks take query blocks slot, slot + S, ... (S blocks per XCD)
        // and sweep the XCD's ranges with each, all in the same order, so the
        // XCD's blocks work on one range at a time and each block stages a
        // query image once per sweep.
        const int x = blockIdx.x % 8, slot = blockIdx.x / 8, nslot = gridDim.x / 8;
        const int r0 = (int)((int64_t)nranges * x / 8), r1 = (int)((int64_t)nranges * (x + 1) / 8);
        for (int64_t qb = slot; qb < nqb; qb += nslot)
            for (int rg = r0; rg < r1; rg++) {
                __syncthreads();   // the previous item's LDS readers are done
                process(qb, rg, rg == r0);
            }
    }
    block_add_u64(nver, verified);
}

constexpr int TPB = 256;   // pivots per block of the pivot table kernel (a multiple of 256)

// Distance table of the L2 radius features, D[q][p] = d(query q, pivot row
// p), from one BF16 hi x hi MFMA product per step: d^2 carries an error of
// order 2^-8 |q| |p|, immaterial to the features (log-scaled distances of the
// nearest pivots), which training and inference both take from this table.
// Block = 64 queries (the hi half of their LDS image, 42 KB at dk = 320) x
// TPB pivots: 8 waves, each 64 x 32 per tile; two blocks per CU (at most 128
// registers).
__global__ void __launch_bounds__(512, 4) k_pivot_table_bf16(const uint16_t *__restrict__ qimg, int dk,
                                                          const int32_t *__restrict__ qpos,
                                                          const float *__restrict__ xn2, int64_t nq,
                                                          const uint16_t *__restrict__ Ph,
                                                          const float *__restrict__ pn2, int np, float *D)
{
    extern __shared__ __attribute__((aligned(16))) uint16_t smem[];
    __shared__ float s_qn[QB];
    const int pitch = dk + 8;
    uint16_t *Qh = smem;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int64_t npb = (np + TPB - 1) / TPB;
    const int64_t qb = blockIdx.x / npb, pb = blockIdx.x % npb;
    const int nk = dk / 32, kl = 8 * (lane >> 4);
    if (tid < QB) {
        const int64_t q = qb * QB + tid;
        s_qn[tid] = q < nq ? xn2[qpos[q]] : 0.0f;
    }
    {
        // the image's hi rows, every load in flight before the LDS stores
        const uint4 *src = (const uint4 *)(qimg + qb * 2 * QB * pitch);
        uint4 *dst = (uint4 *)smem;
        const int nimg = QB * pitch / 8;
        for (int c0 = tid; c0 < nimg; c0 += 4 * blockDim.x) {
            uint4 v[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int c = c0 + u * blockDim.x;
                v[u] = c < nimg ? src[c] : make_uint4(0, 0, 0, 0);
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int c = c0 + u * blockDim.x;
                if (c < nimg) dst[c] = v[u];
            }
        }
    }
    __syncthreads();
    for (int t = 0; t < TPB / 256; t++) {
        const int64_t p0 = pb * TPB + (t * 8 + w) * 32;
        if (p0 >= np) break;
        f32x4 acc[4][2];
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
        const uint16_t *hb = Ph + (p0 / 32 * 2 * nk * 64 + lane) * 8;
        const int64_t blk2 = (int64_t)nk * 512;
        // B fragments two K steps ahead of the MFMAs that use them (three
        // register buffers in rotation); loads past the last step re-read it
        auto ldk = [&](int k, uint4 &h0, uint4 &h1) {
            const int64_t o = (int64_t)min(k, nk - 1) * 512;
            h0 = *(const uint4 *)(hb + o);
            h1 = *(const uint4 *)(hb + o + blk2);
        };
        auto mma = [&](int k, const uint4 &h0, const uint4 &h1) {
            bf16x8 ah[4];
#pragma unroll
            for (int a = 0; a < 4; a++) ah[a] = *(const bf16x8 *)(Qh + (16 * a + (lane & 15)) * pitch + 32 * k + kl);
            const bf16x8 vh[2] = {__builtin_bit_cast(bf16x8, h0), __builtin_bit_cast(bf16x8, h1)};
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 2; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], vh[b], acc[a][b], 0, 0, 0);
            __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);
            __builtin_amdgcn_sched_group_barrier(0x008, 8, 0);
        };
        uint4 xh0, xh1, yh0, yh1, zh0, zh1;
        ldk(0, xh0, xh1);
        ldk(1, yh0, yh1);
        for (int k = 0; k < nk; k += 3) {
            ldk(k + 2, zh0, zh1);
            __builtin_amdgcn_sched_barrier(0);
            mma(k, xh0, xh1);
            __builtin_amdgcn_sched_barrier(0);
            if (k + 1 >= nk) break;
            ldk(k + 3, xh0, xh1);
            __builtin_amdgcn_sched_barrier(0);
            mma(k + 1, yh0, yh1);
            __builtin_amdgcn_sched_barrier(0);
            if (k + 2 >= nk) break;
            ldk(k + 4, yh0, yh1);
            __builtin_amdgcn_sched_barrier(0);
            mma(k + 2, zh0, zh1);
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int b = 0; b < 2; b++) {
            const int64_t p = p0 + 16 * b + (lane & 15);
            if (p >= np) continue;
            const float pn = pn2[p];
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int i = 16 * a + 4 * (lane >> 4) + r;
                    const int64_t q = qb * QB + i;
                    if (q < nq) D[q * np + p] = sqrtf(fmaxf(s_qn[i] + pn - 2.0f * acc[a][b][r], 0.0f));
                }
        }
    }
}

__global__ void k_mask_table(const int32_t *rq, const int32_t *rn, const uint32_t *rm, int64_t nrec,
                             int64_t nP, int32_t pmin, uint32_t *maskT)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nrec;
         i += (int64_t)gridDim.x * blockDim.x)
        maskT[(int64_t)rq[i] * nP + (rn[i] - pmin)] = rm[i];
}

__global__ void k_par_of(const int32_t *nlid, const int32_t *nsize, const uint8_t *nflag,
                         int64_t p0, int64_t np, int32_t *par_of)
{
    // one block per parent node of the level
    for (int64_t b = blockIdx.x; b < np; b += gridDim.x) {
        const int64_t P = p0 + b;
        if ((nflag[P] & 3) != 1) continue;
        const int32_t lid = nlid[P], sz = nsize[P];
        for (int32_t j = threadIdx.x; j < sz; j += blockDim.x) par_of[lid + j] = (int32_t)P;
    }
}

// Dense distance table D[q][c] = d(query q, pivot of node c0 + c) for every
// query of the batch and every split node of one level (FP32, same error
// model as the sparse kernel).  Tile: 64 queries x 128 pivots per block.
template <int M>
__global__ void __launch_bounds__(256) k_dense_pivot(DevTree t, const int32_t *qpos, int64_t nq,
                                                     int64_t c0, int64_t ncol, const int32_t *pivpos, float *D)
{
    __shared__ __attribute__((aligned(16))) float Qs[DKF][QB + 4];
    __shared__ __attribute__((aligned(16))) float Ps[DKF][RT + 4];
    __shared__ int32_t s_qp[QB], s_pp[RT];
    const int tid = threadIdx.x;
    const int64_t ncb = (ncol + RT - 1) / RT;
    const int64_t qb = blockIdx.x / ncb, cb = blockIdx.x % ncb;
    const int dp = t.dp;
    if (tid < QB) {
        int64_t q = qb * QB + tid;
        s_qp[tid] = q < nq ? qpos[q] : -1;
    }
    if (tid < RT) {
        int64_t c = cb * RT + tid;
        s_pp[tid] = c < ncol ? (pivpos ? pivpos[c] : t.npiv[c0 + c]) : -1;
    }
    __syncthreads();
    const int ty = tid >> 4, tx = tid & 15;
    float acc[4][8];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 8; b++) acc[a][b] = 0.0f;
    for (int k0 = 0; k0 < dp; k0 += DKF) {
        const int kw = min(DKF, dp - k0);
        // queries: 64 rows x 32 dims (8 per thread); pivots: 128 rows x 32 dims (16 per thread)
        {
            const int r = tid >> 2, kk = (tid & 3) * 8;
            float v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            if (s_qp[r] >= 0)
                for (int e = 0; e < 8; e++)
                    if (kk + e < kw) v[e] = t.X[(int64_t)s_qp[r] * dp + k0 + kk + e];
            for (int e = 0; e < 8; e++) Qs[kk + e][r] = v[e];
        }
        {
            const int r = tid >> 1, kk = (tid & 1) * 16;
            float v[16];
            for (int e = 0; e < 16; e++) v[e] = 0.0f;
            if (s_pp[r] >= 0)
                for (int e = 0; e < 16; e++)
                    if (kk + e < kw) v[e] = t.X[(int64_t)s_pp[r] * dp + k0 + kk + e];
            for (int e = 0; e < 16; e++) Ps[kk + e][r] = v[e];
        }
        __syncthreads();
        for (int kk = 0; kk < kw; kk++) {
            const float4 q4 = *(const float4 *)&Qs[kk][ty * 4];
            const float4 p0 = *(const float4 *)&Ps[kk][tx * 8];
            const float4 p1 = *(const float4 *)&Ps[kk][tx * 8 + 4];
            const float qa[4] = {q4.x, q4.y, q4.z, q4.w};
            const float pa[8] = {p0.x, p0.y, p0.z, p0.w, p1.x, p1.y, p1.z, p1.w};
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 8; b++) acc[a][b] = Acc<M>::step(acc[a][b], qa[a], pa[b]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; a++) {
        const int64_t q = qb * QB + ty * 4 + a;
        if (q >= nq) continue;
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const int64_t c = cb * RT + tx * 8 + b;
            if (c < ncol)
                D[q * ncol + c] = s_pp[tx * 8 + b] >= 0 ? tile_fin<M>(t, acc[a][b], s_qp[ty * 4 + a], s_pp[tx * 8 + b])
                                                       : INFINITY;
        }
    }
}

__global__ void k_gather_dq(const float *D, int64_t ncol, int64_t c0, const int32_t *fq,
                            const int32_t *fn, int64_t nf, float *dq)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nf;
         i += (int64_t)gridDim.x * blockDim.x)
        dq[i] = D[(int64_t)fq[i] * ncol + (fn[i] - c0)];
}

__global__ void k_roff(const int32_t *rec_q, int64_t nrec, int64_t nq, int64_t *roff)
{
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q <= nq;
         q += (int64_t)gridDim.x * blockDim.x) {
        int64_t lo = 0, hi = nrec;
        while (lo < hi) {
            int64_t mid = (lo + hi) >> 1;
            if (rec_q[mid] < q) lo = mid + 1;
            else hi = mid;
        }
        roff[q] = lo;
    }
}

__global__ void k_rec_keys(const int32_t *rq, const int32_t *rn, int64_t n, uint64_t *k, int32_t *v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        k[i] = ((uint64_t)(uint32_t)rq[i] << 32) | (uint32_t)(rn[i] + 1);
        v[i] = (int32_t)i;
    }
}

__global__ void k_rec_gather(const int32_t *idx, int64_t n, const int32_t *rq, const int32_t *rn,
                             const uint32_t *rm, int32_t *gq, int32_t *gn, uint32_t *gm)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        int32_t s = idx[i];
        gq[i] = rq[s];
        gn[i] = rn[s];
        gm[i] = rm[s];
    }
}

unsigned grid_for(int64_t n, int block)
{
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (unsigned)g;
}

}  // namespace

bool v2_supported(const gts_index *I, int mode)
{
    // the record-merge and dense kernels cover the BF16 MFMA kNN form
    return mode == V2_KNN && I->verify_form == GTS_VERIFY_BF16X3 && I->metric == GTS_METRIC_L2 &&
           I->dp <= 512;
}

// Records sorted by (query, parent) and per-query offsets.
V2Recs v2_prepare(gts_index *I, Search &S)
{
    hipStream_t st = I->st;
    V2Recs r;
    r.n = S.nrec;
    r.rec_q = I->rec_q.p;
    r.rec_n = I->rec_n.p;
    r.rec_m = I->rec_m.p;
    if (S.rec_levels > 1 && S.nrec > 1) {
        // records from several levels: sort by (query, parent) in the
        // index's grow-only buffers
        uint64_t *k0 = I->rs_k0.ensure(S.nrec), *k1 = I->rs_k1.ensure(S.nrec);
        int32_t *v0 = I->rs_v0.ensure(S.nrec), *v1 = I->rs_v1.ensure(S.nrec);
        k_rec_keys<<<grid_for(S.nrec, 256), 256, 0, st>>>(I->rec_q.p, I->rec_n.p, S.nrec, k0, v0);
        int qb = 1;
        while (qb < 31 && ((int64_t)1 << qb) < S.nq) qb++;
        uint64_t *ks;
        int32_t *vs;
        radix_sort_u64(k0, v0, k1, v1, S.nrec, 32 + qb, nullptr, 0, st, &ks, &vs);
        int32_t *gq = I->grp_q.ensure(S.nrec), *gn = I->grp_n.ensure(S.nrec);
        uint32_t *gm = I->grp_m.ensure(S.nrec);
        k_rec_gather<<<grid_for(S.nrec, 256), 256, 0, st>>>(vs, S.nrec, I->rec_q.p, I->rec_n.p, I->rec_m.p, gq, gn, gm);
        GTS_HIP(hipGetLastError());
        r.rec_q = gq;
        r.rec_n = gn;
        r.rec_m = gm;
    }
    r.roff = I->tile_off.ensure(S.nq + 1);
    k_roff<<<grid_for(S.nq + 1, 256), 256, 0, st>>>(r.rec_q, S.nrec, S.nq, r.roff);
    GTS_HIP(hipGetLastError());
    return r;
}

// One launch of the dense kernel with a schedule: row-range size (rows per
// block item, a multiple of DR) and grid (one block per item, XCD-aware) or
// persistent (one block per CU sweeping its XCD's ranges).
static void dense_kernel_launch(gts_index *I, const Search &S, const VerifyArgs &va, const uint32_t *maskT,
                                int64_t nP, int32_t pmin, const int32_t *par_of, const uint16_t *qimg,
                                DenseSched sc)
{
    hipStream_t st = I->st;
    DevTree t = I->dev();
    const int64_t nqb = (S.nq + QB - 1) / QB;
    const BfErr be = bf_err_model(I->dim, I->dk, 3);
    const int64_t nchunk = (I->n + DR - 1) / DR;
    const int nranges = (int)std::max<int64_t>(1, std::min<int64_t>((I->n + sc.rows - 1) / sc.rows, nchunk));
    // LDS: the query block (hi, lo) and 12 bytes per row of the largest range
    const int64_t max_rows = (nchunk + nranges - 1) / nranges * DR;
    const size_t lds = (size_t)2 * QB * (I->dk + 8) * 2 + (size_t)max_rows * 12;
    if (lds > 160 * 1024) fail(GTS_E_CAPACITY, "dense verification tile exceeds LDS");
    int ncu = 0;
    GTS_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, I->device));
    const int64_t grid = sc.persist ? std::max(8, ncu / 8 * 8) : (nqb * nranges + 7) / 8 * 8;
    if (grid > (int64_t)INT32_MAX) fail(GTS_E_CAPACITY, "too many verification blocks");
    auto go = [&](auto kern) {
        (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        kern<<<(unsigned)grid, DW * 64, lds, st>>>(t, I->Xhi, I->Xlo, I->xn2, I->dk, maskT, nP, pmin, par_of,
                                                  S.nq, nqb, nranges, qimg, S.qpos, S.bnd, be, va.cand,
                                                  va.ccount, va.cap, (unsigned long long *)va.verified);
    };
    if (sc.persist) go(k_verify_dense_bf16<true>);
    else go(k_verify_dense_bf16<false>);
    GTS_HIP(hipGetLastError());
    if (S.nq >= I->last_sched_nq) {   // report the schedule of the call's largest launch
        I->last_sched = sc;
        I->last_sched_nq = S.nq;
    }
}

// The dense kernel over all rows for the batch: leaf-mask table maskT[q][P -
// pmin] with the parent of each row in par_of, or no table (every row).
//
static uint64_t sched_key(const gts_index *I, int64_t nqb, bool masked)
{
    int qlog = 0;
    while (((int64_t)1 << qlog) < nqb) qlog++;
    return ((uint64_t)qlog << 56) ^ ((uint64_t)I->dk << 40) ^ ((uint64_t)masked << 39) ^ (uint64_t)I->n;
}

// Schedule autotuner: the first launch for a shape (query blocks rounded to a
// power of two, padded dimension, rows, leaf-mask or not) times every
// candidate schedule on the batch itself -- each run clears the candidate
// counters first, so every run computes the same candidates -- keeps the
// fastest in the index and reruns it for the result.  Later launches of the
// shape use the stored schedule.
static void dense_launch(gts_index *I, const Search &S, const VerifyArgs &va, const uint32_t *maskT,
                         int64_t nP, int32_t pmin, const int32_t *par_of)
{
    hipStream_t st = I->st;
    const int64_t nqb = (S.nq + QB - 1) / QB;
    const uint16_t *qimg = stage_query_image(I, S.qpos, S.nq);
    // candidates: rows per range from 2.5 MB of BF16 hi/lo rows (fits an XCD's
    // 4 MB L2 with room), twice that, and the most the LDS holds; grid, or
    // persistent when there are enough query blocks to occupy every CU.
    // (Measured at 10,000 queries x 1M x 300: fewer, larger ranges win because
    // each block restages its 84 KB query image; smaller ranges lose up to 3x.)
    const int64_t base = std::min<int64_t>(4096, std::max<int64_t>(DR, (int64_t)(2.5e6 / (4.0 * I->dk)) / DR * DR));
    const int64_t lds_rows = ((int64_t)160 * 1024 - 2048 - (int64_t)2 * QB * (I->dk + 8) * 2) / 12 / DR * DR - DR;
    std::vector<DenseSched> cands;
    for (int64_t rows : {base, 2 * base, lds_rows}) {
        if (rows < DR || rows > lds_rows) continue;
        bool dup = false;
        for (const DenseSched &c : cands) dup |= c.rows == rows;
        if (dup) continue;
        cands.push_back(DenseSched{false, (int)rows});
        if (nqb >= 64) cands.push_back(DenseSched{true, (int)rows});
    }
    if (cands.empty()) cands.push_back(DenseSched{false, (int)std::min<int64_t>(base, std::max<int64_t>(DR, lds_rows))});
    if (I->forced_sched.rows > 0) {
        dense_kernel_launch(I, S, va, maskT, nP, pmin, par_of, qimg, I->forced_sched);
        return;
    }
    const uint64_t key = sched_key(I, nqb, maskT != nullptr);
    auto it = I->dense_sched.find(key);
    if (it != I->dense_sched.end() || !I->autotune || cands.size() == 1) {
        const DenseSched use = it != I->dense_sched.end() ? it->second : cands[0];
        if (I->autotune) I->dense_sched[key] = use;
        dense_kernel_launch(I, S, va, maskT, nP, pmin, par_of, qimg, use);
        return;
    }
    auto reset = [&]() {
        fill_i32(va.ccount, 0, S.nq, st);
        fill_i32((int32_t *)va.verified, 0, 2, st);   // the 64-bit pair counter
    };
    hipEvent_t e0, e1;
    GTS_HIP(hipEventCreate(&e0));
    GTS_HIP(hipEventCreate(&e1));
    float best = INFINITY;
    DenseSched pick = cands[0];
    for (const DenseSched &c : cands) {
        reset();
        GTS_HIP(hipEventRecord(e0, st));
        dense_kernel_launch(I, S, va, maskT, nP, pmin, par_of, qimg, c);
        GTS_HIP(hipEventRecord(e1, st));
        GTS_HIP(hipEventSynchronize(e1));
        float ms = 0;
        GTS_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
            best = ms;
            pick = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    I->dense_sched[key] = pick;
    reset();
    dense_kernel_launch(I, S, va, maskT, nP, pmin, par_of, qimg, pick);
}

uint16_t *stage_query_image(gts_index *I, const int32_t *qpos, int64_t nq)
{
    const int64_t nqb = (nq + QB - 1) / QB;
    uint16_t *qimg = I->qimg.ensure(nqb * 2 * QB * (I->dk + 8));
    if (qpos) {
        k_stage_queries<<<grid_for(nqb * QB * ((I->dk + 8) / 8), 256), 256, 0, I->st>>>(I->X, I->dim, I->dp, I->dk,
                                                                                       qpos, nq, nqb, qimg);
        GTS_HIP(hipGetLastError());
    }
    return qimg;
}

bool dense_schedule_known(const gts_index *I, int64_t nq)
{
    if (!I->Xhi) return false;
    if (I->forced_sched.rows > 0 || !I->autotune) return true;
    return I->dense_sched.count(sched_key(I, (nq + QB - 1) / QB, false)) > 0;
}

void dense_all_launch(gts_index *I, const Search &S, const VerifyArgs &va)
{
    ensure_bf16_split(I);
    dense_launch(I, S, va, nullptr, 0, 0, nullptr);
}

void v2_launch(gts_index *I, Search &S, const V2Recs &r, int mode, const VerifyArgs &va)
{
    (void)mode;
    hipStream_t st = I->st;
    DevTree t = I->dev();
    ensure_bf16_split(I);
    const int64_t nqb = (S.nq + QB - 1) / QB;
    const size_t lds = (size_t)2 * QB * (I->dk + 8) * 2;
    BfErr be = bf_err_model(I->dim, I->dk, 3);
    // dense regime: one record level whose (query, parent) pairs mostly survived
    const int32_t pmin = S.rec_pmin;
    const int64_t nP = (int64_t)S.rec_pmax - S.rec_pmin + 1;
    if (S.rec_levels == 1 && pmin >= 0 && S.nrec * 2 >= S.nq * nP && (double)S.nq * nP * 4 <= 16e9) {
        uint32_t *maskT = (uint32_t *)I->dense.ensure(S.nq * nP);
        GTS_HIP(hipMemsetAsync(maskT, 0, S.nq * nP * 4, st));
        k_mask_table<<<grid_for(S.nrec, 256), 256, 0, st>>>(r.rec_q, r.rec_n, r.rec_m, S.nrec, nP, pmin, maskT);
        if (I->par_level != pmin) {
            int32_t *po = I->par_of.ensure(I->n);
            fill_i32(po, -1, I->n, st);
            k_par_of<<<(unsigned)std::min<int64_t>(nP, 65536), 256, 0, st>>>(I->nlid, I->nsize, I->nflag, pmin, nP, po);
            I->par_level = pmin;
        }
        GTS_HIP(hipGetLastError());
        dense_launch(I, S, va, maskT, nP, pmin, I->par_of.p);
        return;
    }
    // sparse regime: 64-way merge of the queries' record lists
    const int32_t pspan = S.rec_pmax - S.rec_pmin + 1;
    int nranges = (int)std::min<int64_t>(std::max<int64_t>(1, (4096 + nqb - 1) / nqb), 256);
    nranges = std::min(nranges, (int)pspan);
    const int64_t grid = nqb * nranges;
    if (grid > (int64_t)INT32_MAX) fail(GTS_E_CAPACITY, "too many verification blocks");
    (void)hipFuncSetAttribute((const void *)k_verify2_bf16, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    k_verify2_bf16<<<(unsigned)grid, 256, lds, st>>>(t, I->Xhi, I->Xlo, I->xn2, I->dk, r.rec_n, r.rec_m, r.roff,
                                                    S.nq, nqb, pmin, pspan, nranges, S.qpos, S.bnd, be, va.cand,
                                                    va.ccount, va.cap, (unsigned long long *)va.verified);
    GTS_HIP(hipGetLastError());
}

}  // namespace gts

namespace gts {

// d(q, pivot) for the frontier pairs of one level through a dense table when
// the frontier covers a large share of (query, node) pairs.
bool dense_pivot_dist(gts_index *I, const int32_t *qpos, int64_t nq, int lvl, const int32_t *fq,
                      const int32_t *fn, int64_t nf, float *dq)
{
    if (I->metric == GTS_METRIC_EDIT || I->dim < 16) return false;
    const int64_t c0 = I->level_start[lvl], ncol = I->level_start[lvl + 1] - c0;
    if (nf * 4 < nq * ncol) return false;                    // sparse frontier
    if ((double)nq * ncol * 4 > 8.0e9) return false;         // table too large
    if (nq * ncol * 4 > budget_room(I) / 2) return false;    // or above half the memory budget's room
    hipStream_t st = I->st;
    float *D = I->dense.ensure(nq * ncol);
    DevTree t = I->dev();
    const int64_t grid = ((nq + QB - 1) / QB) * ((ncol + RT - 1) / RT);
    GTS_METRIC_SWITCH(I->metric, {
        k_dense_pivot<M_><<<(unsigned)grid, 256, 0, st>>>(t, qpos, nq, c0, ncol, nullptr, D);
    });
    k_gather_dq<<<grid_for(nf, 256), 256, 0, st>>>(D, ncol, c0, fq, fn, nf, dq);
    GTS_HIP(hipGetLastError());
    return true;
}

}  // namespace gts

namespace gts {

// D[q][p] = d(query q, object at tree position piv[p]) for a batch of queries
// and an explicit pivot list (FP32 tiles of 64 queries x 128 pivots).
void pivot_table_bf16(gts_index *I, const int32_t *qpos, int64_t nq, const uint16_t *ph, const float *pn2,
                      int64_t np, float *D)
{
    const uint16_t *qimg = stage_query_image(I, qpos, nq);
    const size_t lds = (size_t)QB * (I->dk + 8) * 2;
    (void)hipFuncSetAttribute((const void *)k_pivot_table_bf16, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    const int64_t grid = ((nq + QB - 1) / QB) * ((np + TPB - 1) / TPB);
    k_pivot_table_bf16<<<(unsigned)grid, 512, lds, I->st>>>(qimg, I->dk, qpos, I->xn2, nq, ph, pn2, (int)np, D);
    GTS_HIP(hipGetLastError());
}

void pivot_table(gts_index *I, const int32_t *qpos, int64_t nq, const int32_t *piv, int64_t np, float *D)
{
    if (I->metric == GTS_METRIC_EDIT) fail(GTS_E_ARG, "pivot tables cover vector metrics");
    const int64_t grid = ((nq + QB - 1) / QB) * ((np + RT - 1) / RT);
    DevTree t = I->dev();
    GTS_METRIC_SWITCH(I->metric, {
        k_dense_pivot<M_><<<(unsigned)grid, 256, 0, I->st>>>(t, qpos, nq, 0, np, piv, D);
    });
    GTS_HIP(hipGetLastError());
}

}  // namespace gts
