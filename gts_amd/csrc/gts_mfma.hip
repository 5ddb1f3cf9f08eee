// L2 verification on the matrix cores (v_mfma_f32_16x16x32_bf16, gfx950).
//
// d^2(q, x) = |q|^2 + |x|^2 - 2 q.x with each operand split into BF16 high
// and low parts: h = bf16(v), l = bf16(v - h), v = h + l + r with
// |v - h| <= u8 |v|, |l| <= u8 (1 + u8) |v|, |r| <= u8^2 |v| (u8 = 2^-8).
// Products of BF16 values are exact in FP32; the kernels accumulate them in
// FP32 (u = 2^-24, K = padded dimension).  An accumulation of N products
// errs by at most gamma_N sum |products| <= 1.02 N u sum |q_i x_i| with
// round-to-nearest additions; the bounds below take twice that, so they also
// hold if the matrix core's internal additions round in one direction.
//   three products, q.x ~ qh.xh + qh.xl + ql.xh: the dropped terms (ql.xl and
//     the residues times the other operand) are at most 3.05 u8^2 |q_i x_i|
//     per component, the accumulation of 3K products adds 2.04 (3K + 8) u;
//   hi only, q.x ~ qh.xh: dropped (qh - q) x + qh (xh - x), at most
//     u8 (2 + u8) |q_i x_i|, accumulation 2.04 (K + 8) u.
// With sum |q_i x_i| <= |q||x| the error of d^2 is bounded per pair by
//   E2 = c1 (|q|^2 + |x|^2) + 2 c2 |q||x|,  c1 = (dim + 8) u (FP32 norms and
// the final additions), c2 the sum of the two terms of the form,
// and |d - d_hat| <= min(sqrt(E2), E2 / d_hat).  A pair is a kNN candidate
// when d_hat - e <= bound; the top-k selection re-ranks every candidate whose
// band reaches the k-th distance in FP64, so result sets stay exact.
//
// Tile: 64 queries x 64 objects per 256-thread block; each wave owns a
// 32 x 32 quadrant = 2 x 2 MFMA blocks of 16 x 16, K = 32 dims per stage,
// 3 MFMAs per block and stage.  Objects are the contiguous positions of runs
// of surviving leaf children of one parent node (as in the FP32 kernel).
#include "gts_dist.h"
#include "gts_index.h"
#include "gts_mfma.h"
#include "gts_wave.h"
#include "gts_search.h"

namespace gts {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int QT = 64, RT = 64, KS = 32;
constexpr int LP = KS + 8;  // LDS pitch in bf16 (80 bytes, 16-byte aligned rows)

unsigned grid_for(int64_t n, int block)
{
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (unsigned)g;
}

// BF16 hi/lo split of the tree-ordered rows into the fragment-major layout
// (rows padded to a multiple of 32 with zeros, the dense kernel's row chunk).
__global__ void k_split_bf16(const float *X, int64_t n, int64_t npad, int dim, int dp, int dk,
                             uint16_t *hi, uint16_t *lo)
{
    const int nks = dk / 32;
    const int64_t total = npad * dk;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t pos = i / dk;
        const int j = (int)(i % dk);
        const float v = (pos < n && j < dim) ? X[pos * dp + j] : 0.0f;
        const uint16_t h = bf16_rn(v);
        const float r = v - __uint_as_float((uint32_t)h << 16);
        const int64_t o = frag_off(pos, j & ~7, nks) + (j & 7);
        hi[o] = h;
        lo[o] = bf16_rn(r);
    }
}

__device__ __forceinline__ int64_t ub_i64(const int64_t *a, int64_t n, int64_t v)
{
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (a[mid] <= v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(256) k_verify_bf16(DevTree t, const uint16_t *__restrict__ Xh,
                                                     const uint16_t *__restrict__ Xl,
                                                     const float *__restrict__ xn2, int dk,
                                                     const int32_t *gq, const uint32_t *gm,
                                                     const int64_t *node_off, const int64_t *tile_off,
                                                     int64_t nslots, const int32_t *qpos,
                                                     const float *bnd, BfErr be, uint64_t *cand,
                                                     int32_t *ccount, int cap,
                                                     unsigned long long *verified)
{
    __shared__ __attribute__((aligned(16))) uint16_t Qh[QT][LP], Ql[QT][LP];
    __shared__ __attribute__((aligned(16))) uint16_t Xsh[RT][LP], Xsl[RT][LP];
    __shared__ int32_t s_q[QT], s_qp[QT];
    __shared__ uint32_t s_m[QT];
    __shared__ float s_b[QT], s_qn[QT];
    __shared__ uint32_t s_union;
    __shared__ int32_t s_cnt[QT];
    __shared__ int32_t s_base[QT];

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int64_t tile = blockIdx.x;
    const int64_t slot = ub_i64(tile_off, nslots + 1, tile) - 1;
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
            s_qn[tid] = xn2[qpos[q]];
            atomicOr(&s_union, gm[rstart + tid]);
        } else {
            s_q[tid] = -1;
            s_qp[tid] = 0;
            s_m[tid] = 0;
            s_b[tid] = -1.0f;
            s_qn[tid] = 0.0f;
        }
    }
    __syncthreads();
    const uint32_t umask = s_union;
    const int nchild = P < 0 ? 1 : t.T;
    const int32_t plid = P < 0 ? 0 : t.nlid[P];
    const int32_t pavg = P < 0 ? 0 : t.nsize[P] / t.T;
    const int qb = (w >> 1) * 32, rb = (w & 1) * 32;
    const int lrow = tid >> 2, lseg = (tid & 3) * 8;   // staging: row, 8-element segment
    unsigned long long nver = 0;

    for (int c = 0; c < nchild; c++) {
        if (!((umask >> c) & 1u)) continue;
        int cend = c;
        while (cend + 1 < nchild && ((umask >> (cend + 1)) & 1u)) cend++;
        int32_t clid, run;
        if (P < 0) {
            clid = 0;
            run = (int32_t)t.n;
        } else {
            int64_t c0 = (int64_t)P * t.T + 1;
            clid = t.nlid[c0 + c];
            run = t.nlid[c0 + cend] + t.nsize[c0 + cend] - clid;
        }
        c = cend;
        for (int r0 = 0; r0 < run; r0 += RT) {
            f32x4 acc[2][2];
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
            const bool qv = lrow < rcount;
            const bool xv = r0 + lrow < run;
            const int64_t qrow = s_qp[lrow];
            const int64_t xrow = clid + r0 + lrow;
            for (int k0 = 0; k0 < dk; k0 += KS) {
                uint4 z = make_uint4(0, 0, 0, 0);
                const int64_t qo = frag_off(qrow, k0 + lseg, dk / 32), xo = frag_off(xrow, k0 + lseg, dk / 32);
                uint4 a0 = qv ? *(const uint4 *)(Xh + qo) : z;
                uint4 a1 = qv ? *(const uint4 *)(Xl + qo) : z;
                uint4 b0 = xv ? *(const uint4 *)(Xh + xo) : z;
                uint4 b1 = xv ? *(const uint4 *)(Xl + xo) : z;
                *(uint4 *)&Qh[lrow][lseg] = a0;
                *(uint4 *)&Ql[lrow][lseg] = a1;
                *(uint4 *)&Xsh[lrow][lseg] = b0;
                *(uint4 *)&Xsl[lrow][lseg] = b1;
                __syncthreads();
                bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
                for (int m = 0; m < 2; m++) {
                    const int qr = qb + 16 * m + (lane & 15);
                    const int xr = rb + 16 * m + (lane & 15);
                    const int kk = 8 * (lane >> 4);
                    ah[m] = *(const bf16x8 *)&Qh[qr][kk];
                    al[m] = *(const bf16x8 *)&Ql[qr][kk];
                    bh[m] = *(const bf16x8 *)&Xsh[xr][kk];
                    bl[m] = *(const bf16x8 *)&Xsl[xr][kk];
                }
#pragma unroll
                for (int a = 0; a < 2; a++)
#pragma unroll
                    for (int b = 0; b < 2; b++) {
                        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
                        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
                        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
                    }
                __syncthreads();
            }
            // epilogue: lane holds dot(q = qb + 16a + 4(lane/16) + r, x = rb + 16b + lane%16)
            if (tid < QT) s_cnt[tid] = 0;
            __syncthreads();
            int slot_local[2][2][4];
            uint32_t hit = 0;
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < 2; b++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        slot_local[a][b][r] = 0;
                        const int i = qb + 16 * a + 4 * (lane >> 4) + r;
                        const int j = rb + 16 * b + (lane & 15);
                        const int32_t pj = clid + r0 + j;
                        const int cj = P < 0 ? 0 : (pavg > 0 ? min((pj - plid) / pavg, t.T - 1) : t.T - 1);
                        const bool valid = i < rcount && r0 + j < run && ((s_m[i] >> cj) & 1u) &&
                                           !hidden(t, pj, s_q[i]);
                        if (!valid) continue;
                        nver++;
                        const float qn = s_qn[i], xn = xn2[pj];
                        float d2 = qn + xn - 2.0f * acc[a][b][r];
                        const float d = sqrtf(fmaxf(d2, 0.0f));
                        const float E2 = be.c1 * (qn + xn) + 2.0f * be.c2 * sqrtf(qn * xn);
                        const float e = fminf(sqrtf(E2), d > 0.0f ? E2 / d : INFINITY);
                        if (s_qp[i] == pj || d - e <= s_b[i]) {
                            hit |= 1u << (a * 8 + b * 4 + r);
                            slot_local[a][b][r] = atomicAdd(&s_cnt[i], 1);
                        }
                    }
            __syncthreads();
            if (tid < QT && s_cnt[tid] > 0) s_base[tid] = atomicAdd(&ccount[s_q[tid]], s_cnt[tid]);
            __syncthreads();
            if (hit) {
#pragma unroll
                for (int a = 0; a < 2; a++)
#pragma unroll
                    for (int b = 0; b < 2; b++)
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                            if (!(hit & (1u << (a * 8 + b * 4 + r)))) continue;
                            const int i = qb + 16 * a + 4 * (lane >> 4) + r;
                            const int j = rb + 16 * b + (lane & 15);
                            const int32_t pj = clid + r0 + j;
                            const float qn = s_qn[i], xn = xn2[pj];
                            float d = sqrtf(fmaxf(qn + xn - 2.0f * acc[a][b][r], 0.0f));
                            if (s_qp[i] == pj) d = 0.0f;
                            const int s = s_base[i] + slot_local[a][b][r];
                            if (s < cap)
                                cand[(int64_t)s_q[i] * cap + s] =
                                    ((uint64_t)__float_as_uint(d) << 32) | (uint32_t)t.perm[pj];
                        }
            }
            __syncthreads();
        }
    }
    block_add_u64(nver, verified);
}

}  // namespace

BfErr bf_err_model(int dim, int dk, int nprod)
{
    const double u = 5.9604644775390625e-08, u8 = 1.0 / 256.0;
    BfErr e;
    e.c1 = (float)((dim + 8) * u);
    if (nprod == 1)
        e.c2 = (float)(u8 * (2.0 + u8) + 2.04 * (dk + 8) * u);
    else
        e.c2 = (float)(3.05 * u8 * u8 + 2.04 * (3 * dk + 8) * u);
    return e;
}

void ensure_bf16_split(gts_index *I)
{
    if (I->Xhi) return;
    if (I->metric != GTS_METRIC_L2) fail(GTS_E_ARG, "the BF16 form verifies L2 only");
    I->dk = ((I->dp + KS - 1) / KS) * KS;
    const int64_t npad = (I->n + 31) / 32 * 32;
    size_t cnt = (size_t)npad * I->dk;
    if (hipMalloc((void **)&I->Xhi, cnt * 2) != hipSuccess || hipMalloc((void **)&I->Xlo, cnt * 2) != hipSuccess)
        fail(GTS_E_OOM, "BF16 split copies");
    k_split_bf16<<<grid_for((int64_t)cnt, 256), 256, 0, I->st>>>(I->X, I->n, npad, I->dim, I->dp, I->dk, I->Xhi,
                                                                 I->Xlo);
    GTS_HIP(hipGetLastError());
    ensure_xn2_max(I);
}

// the largest |x|^2 bounds the per-query error the selection uses for keys
// from the norm-expansion forms (BF16 hi+lo, FP32 MFMA)
void ensure_xn2_max(gts_index *I)
{
    if (I->xn2_max_ok) return;
    std::vector<float> h(I->n);
    GTS_HIP(hipMemcpyAsync(h.data(), I->xn2, I->n * 4, hipMemcpyDeviceToHost, I->st));
    GTS_HIP(hipStreamSynchronize(I->st));
    float mx = 0.f;
    for (float v : h) mx = v > mx ? v : mx;
    I->xn2_max = mx;
    I->xn2_max_ok = true;
}

BfErr fp32_mfma_err_model(int dp)
{
    // |d^2 - d_exact^2| <= c (|q| + |x|)^2: the FP32 dot product of dp terms
    // and the squared norms (at most dp + 4 roundings each), 2% slack
    BfErr e;
    e.c1 = e.c2 = (float)(1.02 * (dp + 4) * 5.9604644775390625e-08);
    return e;
}

void launch_verify_bf16(gts_index *I, const Search &S, const VerifyArgs &va)
{
    DevTree t = I->dev();
    BfErr be = bf_err_model(I->dim, I->dk, 3);
    k_verify_bf16<<<(unsigned)S.ntiles, 256, 0, I->st>>>(
        t, I->Xhi, I->Xlo, I->xn2, I->dk, I->grp_q.p, I->grp_m.p, I->node_off.p, I->tile_off.p,
        S.nslots, S.qpos, S.bnd, be, va.cand, va.ccount, va.cap, (unsigned long long *)va.verified);
    GTS_HIP(hipGetLastError());
}

}  // namespace gts
