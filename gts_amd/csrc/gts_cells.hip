// Cell scan of the learned kNN form.
//
// The GTS tree partitions by distance rings around random pivots; on
// high-dimensional clustered rows those rings mix every cluster, so a kNN
// batch with the learned radius ends up verifying every (query, row) pair (the
// exhaustive plan).  The cell scan keeps the exhaustive plan's exactness and
// its MFMA kernel but skips the pairs that pivot bounds exclude, the same
// triangle-inequality filtering the tree applies, over a partition that suits
// dense tiles:
//
//  * build (once per tree): P pivot rows (a strided sample, then rows far from
//    every sampled pivot); every row joins the cell of its nearest pivot; rows
//    are copied in (cell, distance to pivot) order -- the scan order -- with
//    cells numbered so that cells whose pivots are nearest neighbours are
//    adjacent.  Each 32-row chunk of the scan order records the distance range
//    of its rows to their pivots.
//  * per batch: queries sorted by their own scan row (so a 64-query block
//    covers few cells); the exact FP32 query x pivot distance table; for each
//    query block, the chunks where some query q may have a row x with
//    d(q, x) <= b_q, i.e. where the lower bound
//        max(d(q, c) - max_x d(x, c), min_x d(x, c) - d(q, c))
//    over the chunk's rows of cell c is <= b_q (b_q: the learned radius the
//    dense kernel's gate uses); the dense kernel verifies exactly those
//    (query block, chunk) pairs.  Every pair with d(q, x) <= b_q is verified,
//    so the candidate lists -- and the results -- equal the exhaustive plan's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <numeric>
#include <vector>

#include "gts_cells.h"
#include "gts_dist.h"
#include "gts_index.h"
#include "gts_mfma.h"
#include "gts_search.h"
#include "gts_verify2.h"
#include "gts_wave.h"

typedef float float2_t __attribute__((ext_vector_type(2)));

namespace gts {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int QB = 64;        // queries per block
constexpr int DR = 32;        // rows per chunk
constexpr int CW = 12;        // waves per block of the verification kernel (3 per SIMD)
constexpr int CNB = 3;        // K steps in flight per wave (register buffers)
constexpr int CS = 96;        // most chunks per work item (8 per wave)
constexpr int PCS = 32;       // ints between the verification kernel's per-query counters (128 B)
// Radius levels of the verification kernel's tightening: level t has the
// squared radius b0^2 (1 - t BSPAN / NBIN), t = 0..NBIN (b0: the call's radius)
constexpr int NBIN = 32;
constexpr float BSPAN = 0.75f;
constexpr int BPITCH = 36;    // ints per query's level counts (NBIN + 1, padded to 16 bytes)
constexpr int QSORT = 65536;  // queries per ordered segment
constexpr int64_t QSEG = 65536;   // queries per cell pass (bounds the distance table)
constexpr int64_t ASSIGN_ROWS = 16384;   // rows per assignment table
constexpr int FAR_ROUNDS = 3;            // rounds of far-row pivots of the cell partition

unsigned grid_for(int64_t n, int block)
{
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + block - 1) / block, 1 << 20));
}

__global__ void k_gather_f32(const float *src, const int32_t *idx, int64_t n, float *dst)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[idx[i]];
}

// nearest pivot of each row of a distance table D[r][c] (one wave per row;
// ties to the lower cell).  skip0 >= 0: row r ignores column skip0 + r (the
// pivot x pivot table excludes each pivot itself).
__global__ void k_cell_argmin(const float *__restrict__ D, int64_t nr, int P, int64_t skip0, int32_t *cell,
                              float *dist)
{
    const int lane = threadIdx.x & 63;
    const int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (r >= nr) return;
    float best = INFINITY;
    int bi = INT_MAX;
    for (int c = lane; c < P; c += 64) {
        const float v = D[r * P + c];
        if (skip0 >= 0 && c == skip0 + r) continue;
        if (v < best || (v == best && c < bi)) {
            best = v;
            bi = c;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const float ov = __shfl_xor(best, o, 64);
        const int oi = __shfl_xor(bi, o, 64);
        if (ov < best || (ov == best && oi < bi)) {
            best = ov;
            bi = oi;
        }
    }
    if (lane == 0) {
        cell[r] = bi == INT_MAX ? 0 : bi;
        if (dist) dist[r] = best;
    }
}

// sort keys of the scan order: (rank of the row's cell, distance bits)
__global__ void k_scan_keys(const int32_t *cell, const float *dist, const int32_t *rank, int64_t n, uint64_t *keys,
                            int32_t *vals)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        keys[i] = ((uint64_t)(uint32_t)rank[cell[i]] << 32) | __float_as_uint(dist[i]);
        vals[i] = (int32_t)i;
    }
}

// per scan row: inverse map, object ID, norm, cell and distance
__global__ void k_scan_rows(const uint64_t *keys, const int32_t *spos, int64_t n, const int32_t *perm,
                            const float *xn2, int32_t *pos_scan, int32_t *sid, float *sxn2, int32_t *scell,
                            float *sdist)
{
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += (int64_t)gridDim.x * blockDim.x) {
        const int32_t p = spos[s];
        pos_scan[p] = (int32_t)s;
        sid[s] = perm[p];
        sxn2[s] = xn2[p];
        scell[s] = (int32_t)(keys[s] >> 32);
        sdist[s] = __uint_as_float((uint32_t)keys[s]);
    }
}

// BF16 split rows copied from tree order to scan order (fragment-major both;
// padding rows zero): one 8-element group per thread
__global__ void k_scan_split(const uint16_t *__restrict__ hi, const uint16_t *__restrict__ lo,
                             const int32_t *__restrict__ spos, int64_t n, int64_t npad, int dk, uint16_t *shi,
                             uint16_t *slo)
{
    const int nks = dk / 32, ng = dk / 8;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npad * ng;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = i / ng;
        const int k = (int)(i % ng) * 8;
        const int64_t o = frag_off(s, k, nks);
        uint4 h = make_uint4(0, 0, 0, 0), l = make_uint4(0, 0, 0, 0);
        if (s < n) {
            const int64_t src = frag_off(spos[s], k, nks);
            h = *(const uint4 *)(hi + src);
            l = *(const uint4 *)(lo + src);
        }
        *(uint4 *)(shi + o) = h;
        if (slo) *(uint4 *)(slo + o) = l;
    }
}

// FP8 residual rows of the scan order, one thread per scan row: r = x - hi
// (exact in FP32; hi = bf16(x) as the split copies hold it) scaled by
// 2^shift and converted to FP8 pairs (v_cvt_pk_fp8_f32); the decoded value,
// as the verification kernel forms it (v_cvt_scalef32_pk_bf16_fp8: FP8 pair
// times 2^-shift to a BF16 pair), gives the row's residual error
// eps = |x - hi - lo|_2,
// accumulated in FP64 and rounded up.  |q.(x - hi - lo)| <= |q| eps bounds
// what the FP8 plane adds to the dot product's error, whatever the
// conversion's rounding.  A pair that does not decode to finite values is
// stored as zeros (its residual then counts in eps).
__global__ void k_scan_lo8(const float *__restrict__ X, int dp, int dim, int dk, const int32_t *__restrict__ spos,
                           int64_t n, int64_t npad, int shift, uint8_t *lo8, float *eps)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= npad) return;
    const int nks = dk / 32;
    const float sc = ldexpf(1.0f, shift), ds = ldexpf(1.0f, -shift);
    const bool ok = s < n;
    const float *x = X + (ok ? (int64_t)spos[s] * dp : 0);
    double e2 = 0.0;
    for (int k = 0; k < dk; k += 8) {
        float xv[8], hv[8];
#pragma unroll
        for (int e = 0; e < 8; e++) {
            xv[e] = (ok && k + e < dim) ? x[k + e] : 0.0f;
            hv[e] = __uint_as_float((uint32_t)bf16_rn(xv[e]) << 16);
        }
        uint32_t w[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            int v = 0;
#pragma unroll
            for (int pr = 0; pr < 2; pr++) {
                const int e = 4 * h + 2 * pr;
                const float r0 = (xv[e] - hv[e]) * sc, r1 = (xv[e + 1] - hv[e + 1]) * sc;
                int t = __builtin_amdgcn_cvt_pk_fp8_f32(r0, r1, 0, false);   // low 16 bits
                // decoded exactly as the verification kernel decodes it
                const bf16x2 d = __builtin_amdgcn_cvt_scalef32_pk_bf16_fp8((unsigned)t, ds, false);
                float d0 = (float)d.x, d1 = (float)d.y;
                if (!isfinite(d0) || !isfinite(d1)) {
                    t = 0;
                    d0 = d1 = 0.0f;
                }
                v |= (t & 0xFFFF) << (16 * pr);
                const double a = (double)xv[e] - (double)hv[e] - (double)d0;
                const double b = (double)xv[e + 1] - (double)hv[e + 1] - (double)d1;
                e2 += a * a + b * b;
            }
            w[h] = (uint32_t)v;
        }
        *(uint2 *)(lo8 + frag8_off(s, k, nks)) = make_uint2(w[0], w[1]);
    }
    const double r = sqrt(e2) * (1.0 + 1e-9);
    float f = (float)r;
    if ((double)f < r) f = nextafterf(f, INFINITY);
    eps[s] = f;
}

// The verification kernel's row layout: for chunk c and K step k, the 3 KB
// piece at (c * nks + k) * 3072 holds the step's BF16 hi fragments of rows
// 0-15 (1 KB, lane l's 16 bytes at 16 l), of rows 16-31, then its FP8
// residual fragments (frag8_off order), so that a wave's loads of one step
// fall in one contiguous piece and a chunk's 10-30 KB are read front to back.
// One thread per 16 bytes.
__global__ void k_interleave_rows(const uint16_t *__restrict__ hi, const uint8_t *__restrict__ lo8, int64_t nchunk,
                                  int nks, uint8_t *rows)
{
    const int64_t total = nchunk * nks * 192;   // 16-byte pieces
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t piece = i / 192;           // (chunk, step)
        const int part = (int)(i % 192) / 64, lane = (int)(i % 64);
        const int64_t c = piece / nks, k = piece % nks;
        uint4 v;
        if (part < 2) v = *(const uint4 *)(hi + (((2 * c + part) * nks + k) * 64 + lane) * 8);
        else v = *(const uint4 *)(lo8 + ((c * nks + k) * 64 + lane) * 16);
        *(uint4 *)(rows + i * 16) = v;
    }
}

// per cell: |pivot|^2, its root, and the conservative ends of the rows'
// distance range to the pivot
__global__ void k_cell_pack(const float *pn2, const int2 *cell_rows, const float *sdist, int64_t P, ErrModel em,
                            float4 *pack)
{
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < P; c += (int64_t)gridDim.x * blockDim.x) {
        const int2 cr = cell_rows[c];
        const float lo = cr.x >= 0 ? sdist[cr.x] : 0.0f, hi = cr.x >= 0 ? sdist[cr.y] : 0.0f;
        pack[c] = make_float4(pn2[c], sqrtf(pn2[c]), lo - (em.rel * lo + em.abs), hi + (em.rel * hi + em.abs));
    }
}

// per cell: first and last scan row (rows of a cell are contiguous, in
// distance order)
__global__ void k_cell_ranges(const int32_t *scell, int64_t n, int2 *cell_rows)
{
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += (int64_t)gridDim.x * blockDim.x) {
        const int32_t c = scell[s];
        if (s == 0 || scell[s - 1] != c) cell_rows[c].x = (int32_t)s;
        if (s == n - 1 || scell[s + 1] != c) cell_rows[c].y = (int32_t)s;
    }
}

// Batch order: each segment of QSORT queries ordered by the scan row of the
// query, ties by batch index (so the order is deterministic).  The rank of a
// query is the number of (scan row, index) keys of its segment below its
// own: block (i-block, j-tile) counts the keys of a RTILE-key tile below each
// of 256 queries and adds the count to their ranks (zeroed before).
// RTILE: a multiple of 256 (512 measured slower at 10,000 queries).
constexpr int RTILE = 256;

__global__ void __launch_bounds__(256) k_qrank(const int32_t *qpos, const int32_t *inv, const int32_t *pos_scan,
                                               int64_t nq, int64_t n, int NI, int NJ, int32_t *rank)
{
    __shared__ uint64_t sk[RTILE];
    const int64_t seg = blockIdx.x / (NI * NJ);   // NI i-blocks x NJ j-tiles per segment
    const int ib = (int)(blockIdx.x % (NI * NJ)) / NJ, jt = (int)(blockIdx.x % NJ);
    const int64_t s0 = seg * QSORT;
    const int len = (int)min((int64_t)QSORT, nq - s0);
    if (ib * 256 >= len || jt * RTILE >= len) return;   // uniform per block
    const int tl = min(RTILE, len - jt * RTILE);
    // the tile's keys and the thread's own, every load of a level in flight
    // together (tree position, then scan row)
    constexpr int NU = RTILE / 256;
    const int i = ib * 256 + threadIdx.x;
    int32_t qp[NU + 1], ps[NU + 1];
#pragma unroll
    for (int u = 0; u <= NU; u++) {
        const int idx = u < NU ? jt * RTILE + (int)threadIdx.x + 256 * u : i;
        const bool ok = u < NU ? (int)threadIdx.x + 256 * u < tl : i < len;
        qp[u] = ok ? qpos[s0 + idx] : 0;
    }
    if (inv) {   // qpos holds query IDs: their tree positions (IDs out of range,
                 // reported by k_qplace_fields, clamped)
#pragma unroll
        for (int u = 0; u <= NU; u++) qp[u] = inv[min(max(qp[u], 0), (int32_t)(n - 1))];
    }
#pragma unroll
    for (int u = 0; u <= NU; u++) ps[u] = pos_scan[qp[u]];
#pragma unroll
    for (int u = 0; u < NU; u++) {
        const int t = (int)threadIdx.x + 256 * u;
        if (t < tl) sk[t] = ((uint64_t)(uint32_t)ps[u] << 32) | (uint32_t)(jt * RTILE + t);
    }
    __syncthreads();
    if (i >= len) return;
    const uint64_t mine = ((uint64_t)(uint32_t)ps[NU] << 32) | (uint32_t)i;
    int c = 0;
#pragma unroll 8
    for (int u = 0; u < tl; u++) c += sk[u] < mine;
    if (c) atomicAdd(&rank[s0 + i], c);
}

__global__ void k_qplace(const int32_t *rank, int64_t nq, int32_t *qperm)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += (int64_t)gridDim.x * blockDim.x)
        qperm[i / QSORT * QSORT + rank[i]] = (int32_t)i;
}

__global__ void k_qgather(const int32_t *qperm, int64_t nq, const int32_t *qpos, const int32_t *kq,
                          const float *bnd, const int32_t *orig, int32_t *t, float *tf)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t j = qperm[i];
        t[i] = qpos[j];
        t[nq + i] = kq[j];
        t[2 * nq + i] = orig[j];
        tf[i] = bnd[j];
    }
}

__global__ void k_qscatter(int64_t nq, const int32_t *t, const float *tf, int32_t *qpos, int32_t *kq, float *bnd,
                           int32_t *orig)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += (int64_t)gridDim.x * blockDim.x) {
        qpos[i] = t[i];
        kq[i] = t[nq + i];
        orig[i] = t[2 * nq + i];
        bnd[i] = tf[i];
    }
}

// Batch preparation in the scan order, step 2 (after k_qrank over the query
// IDs): every field at its ordered slot (tree position, k, orig: the batch
// index, bnd: +inf), with the pass's candidate counts, its 64-bit scalars scal[8..14]
// and nbits words of the live-chunk bitmap cleared; each rank is cleared once
// used, so the rank array is all zero for the next batch.
__global__ void k_qplace_fields(int32_t *rank, const int32_t *qids, const int32_t *inv, int64_t n, const int32_t *ks,
                                int kfix, int kcap, int32_t *argerr, int64_t nq, int32_t *qpos, int32_t *kq,
                                int32_t *orig, float *bnd, int32_t *ccount, int32_t *scal8, uint32_t *bits,
                                int64_t nbits)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x, i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t i = i0; i < nq; i += stride) {
        const int64_t r = i / QSORT * QSORT + rank[i];
        rank[i] = 0;
        // IDs outside [0, n) and k outside [1, kcap] (device inputs) flag the
        // call (*argerr, published with the pass) and are clamped
        int32_t q = qids[i], k = ks ? ks[i] : kfix;
        if (q < 0 || q >= n || k < 1 || k > kcap) {
            atomicOr(argerr, 1);
            q = min(max(q, 0), (int32_t)(n - 1));
            k = min(max(k, 1), kcap);
        }
        qpos[r] = inv[q];
        kq[r] = k;
        orig[r] = (int32_t)i;
        bnd[r] = INFINITY;
        ccount[i] = 0;
    }
    for (int64_t i = i0; i < nbits; i += stride) bits[i] = 0u;
    if (i0 < 14) scal8[i0] = 0;
}

// conservative FP32 distance range: |d_fp32 - d| <= rel d + abs
__device__ __forceinline__ float dlo(float d, ErrModel em) { return d - (em.rel * d + em.abs); }
__device__ __forceinline__ float dhi(float d, ErrModel em) { return d + (em.rel * d + em.abs); }

constexpr int BW = 8;          // waves per block of the bounds kernel
constexpr int BPT = BW * 32;   // pivots per pass of a bounds block
constexpr int BPASS = 2;       // most passes per block
constexpr int BNB = 4;         // K steps in flight per wave of the bounds kernel

// Live chunks, step 1 (fused distance table and bounds): block = 64 queries
// (the hi half of the dense kernel's LDS image) x bpass * BPT cell pivots.
// Each wave computes 64 x 32 query-pivot dot products with BF16 hi x hi MFMAs
// (the pivot rows' hi copies, fragment-major) and from them the range of the
// exact distance, d2 +- E2 with the error bound of the hi-only product.  A
// lane holds 16 queries of one pivot: it tests its pivot's whole cell against
// those queries and marks the cell's chunks live (a cell of more than 64
// chunks: each chunk, tested with its rows' distance range) in the query
// block's bitmap row, kept in LDS when it fits and merged into the global row
// (cleared beforehand) at the end.  Four waves per SIMD, two blocks per CU:
// one block per CU at 137 registers measured 79 us, two at 128 (40 bytes of
// spills) 57 us, on the clustered 1M x 300 rows with 10k queries.
__global__ void __launch_bounds__(BW * 64, 4) k_cell_bounds(
    const uint16_t *__restrict__ qimg, int dk, const int32_t *__restrict__ qpos, const float *__restrict__ xn2,
    const float *__restrict__ bnd, int64_t nq, int64_t q0, const uint16_t *__restrict__ Ph,
    const float4 *__restrict__ cell_pack, int P, const int2 *__restrict__ cell_rows,
    const float *__restrict__ sdist, int64_t nw, int bpass, bool lds_bits, BfErr be, ErrModel em, uint32_t *bits)
{
    extern __shared__ __attribute__((aligned(16))) uint16_t smem[];
    __shared__ float s_qn[QB], s_b[QB], s_sq[QB];
    const int pitch = dk + 8;
    uint16_t *Qh = smem;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int64_t npb = (P + BPT * bpass - 1) / (BPT * bpass);
    const int64_t qb = blockIdx.x / npb, pb = blockIdx.x % npb;
    const int nk = dk / 32, kl = 8 * (lane >> 4), g = lane >> 4;
    if (tid < QB) {
        const int64_t q = q0 + qb * QB + tid;
        const bool ok = q < nq;
        s_qn[tid] = ok ? xn2[qpos[q]] : 0.0f;
        // the radius with a relative margin of 10^-6 (NaN: an inert padding query)
        s_b[tid] = ok ? bnd[q] * 1.000001f + 1e-30f : -INFINITY;
        s_sq[tid] = sqrtf(s_qn[tid]);
    }
    {
        // the hi half of the image, read with every load in flight before the
        // LDS stores
        const uint4 *src = (const uint4 *)(qimg + (q0 / QB + qb) * 2 * QB * pitch);
        uint4 *dst = (uint4 *)smem;
        const int nimg = QB * pitch / 8;
        for (int c0 = tid; c0 < nimg; c0 += 8 * blockDim.x) {
            uint4 v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int c = c0 + u * blockDim.x;
                v[u] = c < nimg ? src[c] : make_uint4(0, 0, 0, 0);
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int c = c0 + u * blockDim.x;
                if (c < nimg) dst[c] = v[u];
            }
        }
    }
    // live marks go to an LDS copy of the bitmap row when it fits (lds_bits),
    // merged into the global row once per word at the end: many lanes mark
    // the same few words, which as global atomics serialise
    uint32_t *g_bits = bits + qb * nw;
    uint32_t *row_bits = lds_bits ? (uint32_t *)(smem + QB * pitch) : g_bits;
    if (lds_bits)
        for (int i = tid; i < nw; i += blockDim.x) row_bits[i] = 0;
    __syncthreads();
    const float c1 = be.c1 + 8.0f * 5.9604645e-08f;
    for (int pass = 0; pass < bpass; pass++) {
        const int64_t p0 = (pb * bpass + pass) * BPT + w * 32;   // the wave's 32 pivots
        if (p0 >= P) break;
        f32x4 acc[4][2];
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
        const uint16_t *hb = Ph + (p0 / 32 * 2 * nk * 64 + lane) * 8;
        const int64_t blk2 = (int64_t)nk * 512;
        // the lane's two cells (pivot norm and its root, conservative distance
        // range of the rows; rows of an empty cell: x = -1), loaded before
        // the MFMAs so that their latency overlaps them
        float4 cpk[2];
        int2 crw[2];
#pragma unroll
        for (int b = 0; b < 2; b++) {
            const int64_t c = p0 + 16 * b + (lane & 15);
            cpk[b] = c < P ? cell_pack[c] : make_float4(0.f, 0.f, 0.f, 0.f);
            crw[b] = c < P ? cell_rows[c] : make_int2(-1, -1);
        }
        // hi x hi products only (the bound's error model covers the dropped
        // lo terms); B fragments BNB - 1 K steps ahead in BNB rotating
        // register buffers (the pivot rows come from L2: their latency, not
        // the MFMAs, sets this loop's pace); loads past the last step re-read it
        int lk = 0;
        auto ldk = [&](uint4 (&B)[2]) {
            const int64_t o = (int64_t)min(lk, nk - 1) * 512;
            B[0] = *(const uint4 *)(hb + o);
            B[1] = *(const uint4 *)(hb + o + blk2);
            lk++;
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
        uint4 kb[BNB][2];
#pragma unroll
        for (int b = 0; b < BNB; b++) ldk(kb[b]);
        __builtin_amdgcn_sched_barrier(0);
        for (int k0 = 0; k0 < nk; k0 += BNB) {
#pragma unroll
            for (int b = 0; b < BNB; b++) {
                if (k0 + b < nk) mma(k0 + b, kb[b][0], kb[b][1]);
                __builtin_amdgcn_sched_barrier(0);
                ldk(kb[b]);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
#pragma unroll
        for (int b = 0; b < 2; b++) {
            const int64_t c = p0 + 16 * b + (lane & 15);   // the lane's pivot = cell
            if (c >= P) continue;
            const int2 cr = crw[b];
            if (cr.x < 0) continue;
            const float cn = cpk[b].x, sc = cpk[b].y, rl = cpk[b].z, rh = cpk[b].w;
            // The exact query-pivot distance d satisfies |d^2 - d2| <= E2.  A
            // row x of the cell may lie within b of the query only if
            // d - max d(x,c) <= b and min d(x,c) - d <= b, which with
            // d in [sqrt(max(d2 - E2, 0)), sqrt(d2 + E2)] holds whenever
            //   b + rh >= 0,  d2 - E2 <= (b + rh)^2   and   (rl <= b  or  d2 + E2 >= (rl - b)^2),
            // compared with squares inflated (deflated) by 10^-5.
            // d^2 - E2 and d^2 + E2 of query i (recomputed where needed, not
            // kept: registers bound this kernel's occupancy)
            auto ring = [&](int a, int r, float &lo2, float &hi2) {
                const int i = 16 * a + 4 * g + r;
                const float qn = s_qn[i];
                const float d2 = qn + cn - 2.0f * acc[a][b][r];
                const float E2 = fmaf(c1, qn + cn, 2.0000025f * be.c2 * s_sq[i] * sc);
                lo2 = d2 - E2;
                hi2 = d2 + E2;
            };
            uint32_t pass_bits = 0;
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    float lo2, hi2;
                    ring(a, r, lo2, hi2);
                    const float bb = s_b[16 * a + 4 * g + r];
                    const float up = bb + rh, dn = rl - bb;
                    if (up >= 0.0f && lo2 <= up * up * 1.00001f && (dn <= 0.0f || hi2 >= dn * dn * 0.99999f))
                        pass_bits |= 1u << (a * 4 + r);
                }
            if (!pass_bits) continue;
            const int ch0 = cr.x / DR, ch1 = cr.y / DR;
            if (ch1 - ch0 < 64) {
                // a cell of at most 64 chunks: every chunk holding its rows (on
                // the clustered 1M x 300 rows the per-chunk ring test below
                // removed 0.4% of the live chunks for a quarter of this
                // kernel's time: one lane walks the cell's chunks with
                // dependent loads while its wave and block wait)
                for (int wd = ch0 >> 5; wd <= ch1 >> 5; wd++) {
                    const int lo = max(ch0, wd * 32) - wd * 32, hi = min(ch1, wd * 32 + 31) - wd * 32;
                    const uint32_t m = (hi == 31 ? 0xFFFFFFFFu : ((1u << (hi + 1)) - 1u)) & ~((1u << lo) - 1u);
                    atomicOr(&row_bits[wd], m);
                }
                continue;
            }
            // a larger cell: each chunk holding its rows, tested with the
            // distance range of those rows (four chunks' loads at a time)
            for (int cb0 = ch0; cb0 <= ch1; cb0 += 4) {
                float rmn[4], rmx[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int ch = min(cb0 + u, ch1);
                    rmn[u] = sdist[max(cr.x, ch * DR)];
                    rmx[u] = sdist[min(cr.y, ch * DR + DR - 1)];
                }
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int ch = cb0 + u;
                    if (ch > ch1) break;
                    const float rmin = dlo(rmn[u], em), rmax = dhi(rmx[u], em);
                    bool live = false;
#pragma unroll
                    for (int a = 0; a < 4; a++)
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                            float lo2, hi2;
                            ring(a, r, lo2, hi2);
                            const float bb = s_b[16 * a + 4 * g + r];
                            const float up = bb + rmax, dn = rmin - bb;
                            live |= ((pass_bits >> (a * 4 + r)) & 1u) && up >= 0.0f && lo2 <= up * up * 1.00001f &&
                                    (dn <= 0.0f || hi2 >= dn * dn * 0.99999f);
                        }
                    if (live) atomicOr(&row_bits[ch >> 5], 1u << (ch & 31));
                }
            }
        }
    }
    if (lds_bits) {
        __syncthreads();
        for (int i = tid; i < nw; i += blockDim.x)
            if (row_bits[i]) atomicOr(&g_bits[i], row_bits[i]);
    }
}

// Live chunks, step 2: one workgroup per query block compacts its bitmap row
// into the chunk list (ascending), count in cnt[qb].
__global__ void __launch_bounds__(256) k_cell_compact(const uint32_t *bits, int64_t nw, int64_t nchunk,
                                                      int32_t *list, int32_t *cnt)
{
    __shared__ int32_t s_w[4];   // 256 threads: four waves
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t qb = blockIdx.x;
    const uint32_t *row = bits + qb * nw;
    const int64_t W = (nw + 255) / 256;
    const int64_t w0 = min(nw, tid * W), w1 = min(nw, w0 + W);
    int mine = 0;
    for (int64_t i = w0; i < w1; i++) mine += __popc(row[i]);
    // exclusive prefix over the block: wave scans, then the four wave totals
    int v = mine;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    if (lane == 63) s_w[wv] = v;
    __syncthreads();
    int at = v - mine, total = 0;
#pragma unroll
    for (int x = 0; x < 4; x++) {
        if (x < wv) at += s_w[x];
        total += s_w[x];
    }
    int32_t *out = list + qb * nchunk;
    for (int64_t i = w0; i < w1; i++)
        for (uint32_t m = row[i]; m; m &= m - 1) out[at++] = (int32_t)(i * 32 + __builtin_ctz(m));
    if (tid == 0) cnt[qb] = total;
}

// The verification kernel's candidate counts (one 128-byte line per query)
// added to ccount[q0 + i]; the padded counters and the published radius
// levels beside them are left zero for the next pass.
__global__ void k_cell_counts(int32_t *pcnt, int64_t sq, int64_t q0, int32_t *ccount)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < sq; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = pcnt[i * PCS], l = pcnt[i * PCS + 1];
        if (v) {
            ccount[q0 + i] += v;
            pcnt[i * PCS] = 0;
        }
        if (l) pcnt[i * PCS + 1] = 0;   // the level the verification published
    }
}

// The dense BF16 hi+lo MFMA verification of the exhaustive plan, over the
// scan-order rows of the live chunks.  Blocks are persistent (one per CU):
// the live (query block, chunk) pairs, query block by query block, form one
// sequence (off[qb]: the first pair of qb); blocks b % 8 run on XCD b % 8,
// XCD x takes the x-th eighth of the sequence and its blocks split that range
// into equal contiguous parts (every block streams the same number of
// chunks), processed as items of at most cs chunks of one query block, so a
// block restages its query image only when the query block changes.  Per
// item the block stages the item's chunk list and per-row terms in LDS; each
// wave owns every CW-th chunk of the item and streams its K steps CNB - 1
// ahead of the MFMAs (rotating register buffers): the chunks come from HBM
// with little reuse, so the wave keeps more bytes in flight than the
// tree-order dense kernel does.  A step's three 16-byte buffer loads per lane
// (hi fragments of the chunk's two row halves, FP8 residuals) read one
// contiguous 3 KB piece of Xr (k_interleave_rows), so a chunk is read front
// to back; lanes whose 8 dimensions of the step are all padding read nothing.
// The rows are the MFMAs' A operand and the queries their B operand, so each
// lane ends a chunk with 8 rows of each of 4 queries (cheap per-query
// counts), and the block tightens each query's radius as it verifies pairs
// sure to lie within a smaller one (the candidate path below).  RERUN only
// names the instance that verifies the small rerun batches (a few padded
// query blocks, latency-bound), apart from the batch's own launches.
template <bool RERUN>
__global__ void __launch_bounds__(CW * 64) k_verify_cells_bf16(
    const uint8_t *__restrict__ Xr, float lscale, const float *__restrict__ seps,
    const float *__restrict__ sxn2, const int32_t *__restrict__ sid, int64_t n, int dim, int dk,
    const int32_t *__restrict__ list, int64_t nchunk,
    const int32_t *__restrict__ cnt, int64_t nqb, int cs, int64_t q0, int64_t nq,
    const uint16_t *__restrict__ qimg, const int32_t *__restrict__ qpos, const float *__restrict__ xn2,
    const float *__restrict__ bnd, const int32_t *__restrict__ kq, BfErr be, uint64_t *cand, int32_t *pcnt,
    int cap, unsigned long long *verified, unsigned long long *live, unsigned long long *ncand)
{
    extern __shared__ __attribute__((aligned(16))) uint16_t smem[];
    // per query of the block: filter threshold, |q| (times the row error
    // term), |q|^2, batch index (bits; -1 for an inert slot)
    __shared__ float4 s_qrec[QB];
    // per query: (NBIN / (BSPAN b0^2), slightly enlarged; b0^2), its current
    // radius level, k, and the number of pairs of each level (pairs whose
    // exact distance is sure to lie within that level's radius, counted in
    // the highest such level)
    __shared__ float2 s_qb[QB];
    __shared__ int32_t s_lvl[QB], s_k[QB];
    __shared__ __attribute__((aligned(16))) int32_t s_bin[QB * BPITCH];
    __shared__ int32_t s_ch[CS];
    __shared__ int s_nvq;
    const int pitch = dk + 8;
    uint16_t *Qh = smem, *Ql = smem + QB * pitch;
    float *s_xr = (float *)(smem + 2 * QB * pitch), *s_xn = s_xr + cs * DR;
    int32_t *s_id = (int32_t *)(s_xn + cs * DR);
    int32_t *off = s_id + cs * DR;   // nqb + 1 entries; off[qb]: the first pair of query block qb
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const float c1 = be.c1 + 8.0f * 5.9604645e-08f;
    const int nk = dk / 32;
    const int kl = 8 * (lane >> 4);
    unsigned long long nver = 0, ncnd = 0;
    // the work sequence's offsets: prefix sums of cnt (wave 0, 64 at a time)
    for (int64_t i = tid; i < nqb; i += blockDim.x) off[i + 1] = cnt[i];
    __syncthreads();
    if (w == 0) {
        int carry = 0;
        for (int64_t b0 = 0; b0 < nqb; b0 += 64) {
            const bool ok = b0 + lane < nqb;
            int v = ok ? off[b0 + lane + 1] : 0;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int u = __shfl_up(v, o, 64);
                if (lane >= o) v += u;
            }
            if (ok) off[b0 + lane + 1] = carry + v;
            carry += __shfl(v, 63, 64);
        }
        if (lane == 0) off[0] = 0;
    }
    __syncthreads();
    const int64_t T = off[nqb];
    if (blockIdx.x == 0 && tid == 0 && live && T) atomicAdd(live, (unsigned long long)T);
    const int x = blockIdx.x % 8, slot = blockIdx.x / 8, nslot = gridDim.x / 8;
    const int64_t xa = T * x / 8, xb = T * (x + 1) / 8;
    const int64_t g0 = xa + (xb - xa) * slot / nslot, g1 = xa + (xb - xa) * (slot + 1) / nslot;
    int64_t cur_qb = -1;
    for (int64_t g = g0; g < g1;) {
        // the item's query block: the last qb with off[qb] <= g (then
        // off[qb + 1] > g, as off[nqb] = T > g)
        int64_t lo = 0, hi = nqb;
        while (hi - lo > 1) {
            const int64_t mid = (lo + hi) >> 1;
            if (off[mid] <= g) lo = mid;
            else hi = mid;
        }
        const int64_t qb = lo, s0c = g - off[qb];
        // chunks of the item, >= 1: to the end of the query block's list or
        // of the block's range, at most cs
        const int m = (int)min(min((int64_t)cs, off[qb + 1] - g), g1 - g);
        g += m;
        const bool new_qb = qb != cur_qb;
        __syncthreads();   // the previous item's LDS readers are done
        if (tid < QB) {
            // the query's level published by any block (pcnt's line of the
            // query, second int): a level other blocks have proven holds for
            // this block's pairs too
            const int gl =
                __hip_atomic_load(&pcnt[(qb * QB + tid) * PCS + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (new_qb) {
                const int64_t q = q0 + qb * QB + tid;
                const bool ok = q < nq && bnd[q] == bnd[q];   // NaN radius: inert padding query of a rerun
                const float qn = ok ? xn2[qpos[q]] : 0.0f;
                const float b = ok ? bnd[q] : -1.0f;
                // level t of a pair with upper bound u^2 on its d^2:
                // floor(NBIN / BSPAN - u^2 NBIN / (BSPAN b0^2)), with the factor
                // enlarged by 2^-14 and the constant lowered by 2^-16 (more than
                // the rounding of the division and of the FMA that forms it), so
                // u^2 <= b0^2 (1 - t BSPAN / NBIN) holds exactly
                const bool lv_ok = ok && b > 0.0f;
                const int l0 = lv_ok ? min(gl, NBIN) : NBIN;   // NBIN: no pair can raise the level
                const float b2 = b * b * (1.0f - (float)(lv_ok ? l0 : 0) * (BSPAN / (float)NBIN));
                // keep a pair iff its exact d^2 may be <= b2 (tree-order kernel's filter)
                s_qrec[tid] = make_float4(ok ? 0.5f * ((1.0f - c1) * qn - b2 * 1.000001f) : INFINITY, sqrtf(qn), qn,
                                          __int_as_float(ok ? (int32_t)q : -1));
                s_qb[tid] = lv_ok ? make_float2((float)NBIN / BSPAN / (b * b) * (1.0f + 6.1035156e-05f), b * b)
                                  : make_float2(0.0f, 0.0f);
                s_lvl[tid] = l0;
                s_k[tid] = ok ? kq[q] : INT_MAX;
                const uint64_t vm = __builtin_amdgcn_ballot_w64(ok);   // wave 0 holds the 64 queries
                if (tid == 0) s_nvq = __popcll(vm);
            } else if (gl > s_lvl[tid] && s_lvl[tid] < NBIN) {
                const int l1 = min(gl, NBIN);
                s_lvl[tid] = l1;
                s_qrec[tid].x = 0.5f * ((1.0f - c1) * s_qrec[tid].z -
                                        s_qb[tid].y * (1.0f - (float)l1 * (BSPAN / (float)NBIN)) * 1.000001f);
            }
        }
        if (tid < m) s_ch[tid] = list[qb * nchunk + s0c + tid];
        if (new_qb) {
            for (int i = tid; i < QB * BPITCH; i += blockDim.x) s_bin[i] = 0;
            const uint4 *src = (const uint4 *)(qimg + (q0 / QB + qb) * 2 * QB * pitch);
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
        cur_qb = qb;
        __syncthreads();   // s_ch visible
        for (int i = tid; i < m * DR; i += blockDim.x) {
            const int64_t row = (int64_t)s_ch[i / DR] * DR + i % DR;
            const bool ok = row < n;
            const float xv = ok ? sxn2[row] : 0.0f;
            s_xr[i] = ok ? fmaf(be.c2, sqrtf(xv), seps[row]) : -1.0f;   // -1: padding row
            s_xn[i] = xv;                        // the candidate path's norm and ID come from LDS
            s_id[i] = ok ? sid[row] : -1;
            if (ok) nver += s_nvq;               // the row's verified pairs
        }
        __syncthreads();
        const int nmy = w < m ? (m - 1 - w) / CW + 1 : 0;
        const int nsteps = nmy * nk;
        if (nsteps == 0) continue;
        f32x4 acc[4][2];
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
        // loader: (chunk ordinal, K step) of the next step to load; steps are
        // loaded in order into CNB rotating register buffers, CNB - 1 steps
        // ahead of the MFMAs; loads past the last step re-read it
        // Loads are buffer loads through a resource over the chunk's nk
        // pieces, issued by every lane on every path: a lane whose 8
        // dimensions of the step are all padding passes an offset beyond the
        // resource and gets zeros without a memory access.  (Loads under a
        // branch leave the compiler unable to count them, and it then waits
        // for every load in flight at the top of each step.)
        int lj = 0, lk = 0;
        const int chunk_bytes = nk * 3072;
        auto ld = [&](uint4 (&B)[3]) {
            const bool past = lj >= nmy;
            const int jj = past ? nmy - 1 : lj, kk = past ? nk - 1 : lk;
            const int64_t c = __builtin_amdgcn_readfirstlane(s_ch[w + jj * CW]);
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                (void *)(Xr + c * chunk_bytes), (short)0, chunk_bytes, 0x00020000);
            const int vo = kk * 32 + kl < dim ? kk * 3072 + lane * 16 : 0x40000000;
            B[0] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, vo, 0, 0));
            B[1] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, vo + 1024, 0, 0));
            B[2] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, vo + 2048, 0, 0));
            const bool lastk = lk == nk - 1;
            lj += lastk ? 1 : 0;
            lk = lastk ? 0 : lk + 1;
        };
        int cj = 0, ck = 0;
        // 8 FP8 residuals (two dwords) -> BF16x8 times 2^-shift, two values
        // per scaled conversion (exact: at most 4 significant bits); the row
        // errors eps were computed from this same decode
        auto dec8 = [&](uint32_t w0, uint32_t w1) {
            auto lo = [&](uint32_t v) {
                return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_scalef32_pk_bf16_fp8(v, lscale, false));
            };
            auto hi = [&](uint32_t v) {
                return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_scalef32_pk_bf16_fp8(v, lscale, true));
            };
            return __builtin_bit_cast(bf16x8, make_uint4(lo(w0), hi(w0), lo(w1), hi(w1)));
        };
        auto step = [&](const uint4 &a0, const uint4 &a1, const uint4 &l8) {
            bf16x8 ah[4], al[4];
#pragma unroll
            for (int a = 0; a < 4; a++) {
                const int qi = 16 * a + (lane & 15);
                ah[a] = *(const bf16x8 *)(Qh + qi * pitch + 32 * ck + kl);
                al[a] = *(const bf16x8 *)(Ql + qi * pitch + 32 * ck + kl);
            }
            const bf16x8 vh[2] = {__builtin_bit_cast(bf16x8, a0), __builtin_bit_cast(bf16x8, a1)};
            const bf16x8 vl[2] = {dec8(l8.x, l8.y), dec8(l8.z, l8.w)};
            // rows as the A operand, queries as B: element r of acc[a][b] is
            // the pair (row 16b + 4(lane/16) + r, query 16a + lane%16)
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 2; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vh[b], ah[a], acc[a][b], 0, 0, 0);
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 2; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vl[b], ah[a], acc[a][b], 0, 0, 0);
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 2; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vh[b], al[a], acc[a][b], 0, 0, 0);
            __builtin_amdgcn_sched_group_barrier(0x100, 8, 0);
            __builtin_amdgcn_sched_group_barrier(0x008, 24, 0);
            if (ck == nk - 1) {
                const int slotc = w + cj * CW;
                const int g = lane >> 4, j = lane & 15;
                // the lane's rows 16b + 4g + r (r = 0..3, consecutive in LDS)
                // and queries 16a + j
                float4 xn4[2], xr4[2];
#pragma unroll
                for (int b = 0; b < 2; b++) {
                    const int li = slotc * DR + 16 * b + 4 * g;
                    xn4[b] = *(const float4 *)&s_xn[li];
                    xr4[b] = *(const float4 *)&s_xr[li];
                }
                const float hx = 0.5f * (1.0f - c1);   // the row term of the filter: hx |x|^2
                float4 qr[4];
#pragma unroll
                for (int a = 0; a < 4; a++) qr[a] = s_qrec[16 * a + j];
                // hb bit a*8 + b*4 + r: the pair (row 16b + 4g + r, query 16a +
                // j) passes the filter (a padding row has xr = -1)
                uint32_t hb = 0;
#pragma unroll
                for (int b = 0; b < 2; b++) {
                    const float ux[4] = {hx * xn4[b].x, hx * xn4[b].y, hx * xn4[b].z, hx * xn4[b].w};
                    const float xr[4] = {xr4[b].x, xr4[b].y, xr4[b].z, xr4[b].w};
#pragma unroll
                    for (int r = 0; r < 4; r++)
#pragma unroll
                        for (int a = 0; a < 4; a++)
                            if (xr[r] >= 0.0f && acc[a][b][r] >= fmaf(-qr[a].y, xr[r], qr[a].x + ux[r]))
                                hb |= 1u << (a * 8 + b * 4 + r);
                }
                // candidate path: a query's 32 rows sit in the four lanes j,
                // j + 16, j + 32, j + 48 (eight each).  The lane's candidate
                // counts of its four queries, one byte each, are scanned over
                // those four lanes; lane (g, j) then reserves the list slots of
                // query 16g + j with one atomic for the chunk, and a candidate's
                // slot is that base, the lower lanes' count and its rank among
                // the lane's own candidates of the query.
                if (__builtin_amdgcn_ballot_w64(hb != 0)) {
                    // Radius tightening: every candidate whose d^2 is sure to
                    // lie within a radius level above its query's current one
                    // (u^2 = d^2 estimate + E2, as the filter's bound the other
                    // way round) is counted in the highest such level; once k
                    // pairs lie at or above a level, the k-th NN distance is
                    // within its radius, and the filter uses that radius for the
                    // rest of the block's work on this query block; the level
                    // is also published (atomic max beside the query's slot
                    // counter) and taken up by every block at its next item of
                    // this query block.  Every pair within the final radius
                    // stays a candidate, so the top-k is unchanged.  A pair is
                    // counted once (each (query, row) pair is verified once per
                    // launch).
                    int lv[4];
                    float lx[4];
#pragma unroll
                    for (int a = 0; a < 4; a++) {
                        lv[a] = s_lvl[16 * a + j];
                        lx[a] = s_qb[16 * a + j].x;
                    }
                    const float c1s = c1 + 16.0f * 5.9604645e-08f;
                    constexpr float LA = (float)NBIN / BSPAN - 1.5258789e-05f;
                    float t2[4][2][4];   // the pairs' d^2 estimates
                    uint32_t sb = 0;     // counted bits, as hb
#pragma unroll
                    for (int a = 0; a < 4; a++)
#pragma unroll
                        for (int b = 0; b < 2; b++)
#pragma unroll
                            for (int r = 0; r < 4; r++) {
                                const float xn = (&xn4[b].x)[r], xr = (&xr4[b].x)[r];
                                const float s2 = qr[a].z + xn;
                                t2[a][b][r] = fmaf(-2.0f, acc[a][b][r], s2);
                                const int bit = a * 8 + b * 4 + r;
                                if ((hb >> bit) & 1u) {
                                    const float ub2 = fmaf(2.0f * qr[a].y, xr, fmaf(c1s, s2, t2[a][b][r]));
                                    const int t = min((int)fmaxf(floorf(fmaf(-ub2, lx[a], LA)), -1.0f), NBIN);
                                    if (t > lv[a]) {
                                        atomicAdd(&s_bin[(16 * a + j) * BPITCH + t], 1);
                                        sb |= 1u << bit;
                                    }
                                }
                            }
                    uint32_t pk = 0, ps = 0;
#pragma unroll
                    for (int a = 0; a < 4; a++) {
                        pk |= (uint32_t)__popc((hb >> (8 * a)) & 0xFFu) << (8 * a);
                        ps |= (uint32_t)__popc((sb >> (8 * a)) & 0xFFu) << (8 * a);
                    }
                    uint32_t x = pk, xs = ps;
                    const uint32_t u1 = (uint32_t)__shfl_up((int)x, 16, 64), v1 = (uint32_t)__shfl_up((int)xs, 16, 64);
                    if (g >= 1) {
                        x += u1;
                        xs += v1;
                    }
                    const uint32_t u2 = (uint32_t)__shfl_up((int)x, 32, 64), v2 = (uint32_t)__shfl_up((int)xs, 32, 64);
                    if (g >= 2) {
                        x += u2;
                        xs += v2;
                    }
                    const uint32_t tot = (uint32_t)__shfl((int)x, 48 + j, 64);
                    const uint32_t stot = (uint32_t)__shfl((int)xs, 48 + j, 64);
                    const uint32_t excl = x - pk;
                    // the query this lane reserves for: L = 16g + j, i.e. a = g
                    // (bytes never carry: at most 32 candidates per query and chunk)
                    float4 qo = qr[0];
#pragma unroll
                    for (int a = 1; a < 4; a++)
                        if (g == a) qo = qr[a];
                    const int mine = (int)((tot >> (8 * g)) & 0xFFu);
                    const int msure = (int)((stot >> (8 * g)) & 0xFFu);
                    ncnd += (unsigned)mine;
                    int32_t id[2][4];
#pragma unroll
                    for (int b = 0; b < 2; b++) {
                        const int4 v = *(const int4 *)&s_id[slotc * DR + 16 * b + 4 * g];
                        id[b][0] = v.x;
                        id[b][1] = v.y;
                        id[b][2] = v.z;
                        id[b][3] = v.w;
                    }
                    int base = 0;
                    // the counters of one query block sit on separate lines (the
                    // block's waves and other blocks reserve for the same queries)
                    if (mine) base = atomicAdd(&pcnt[(int64_t)(__float_as_int(qo.w) - q0) * PCS], mine);
                    int bq[4];
#pragma unroll
                    for (int a = 0; a < 4; a++) bq[a] = __shfl(base, 16 * a + j, 64) + (int)((excl >> (8 * a)) & 0xFFu);
#pragma unroll
                    for (int a = 0; a < 4; a++) {
                        uint64_t *cl = cand + (int64_t)__float_as_int(qr[a].w) * cap;
#pragma unroll
                        for (int b = 0; b < 2; b++)
#pragma unroll
                            for (int r = 0; r < 4; r++) {
                                const int bit = a * 8 + b * 4 + r;
                                if (!((hb >> bit) & 1u)) continue;
                                const int sl = bq[a] + __popc(hb & (((1u << bit) - 1u) & (0xFFu << (8 * a))));
                                // the key: distance bits, ID (v_sqrt_f32 is within 1
                                // ulp, inside the key's error bound, as is the self
                                // pair's rounding residue; the top-k re-ranks its
                                // band in FP64)
                                const float d = __builtin_amdgcn_sqrtf(fmaxf(t2[a][b][r], 0.0f));
                                if (sl < cap) cl[sl] = ((uint64_t)__float_as_uint(d) << 32) | (uint32_t)id[b][r];
                            }
                    }
                    if (msure) {
                        // query L = lane gained counted pairs: its new level is the
                        // highest t with at least k pairs counted at levels >= t
                        const int L = lane, kL = s_k[L];
                        const int4 *bp = (const int4 *)&s_bin[L * BPITCH];
                        int cnt[BPITCH];
#pragma unroll
                        for (int i = 0; i < BPITCH / 4; i++) {
                            const int4 v = bp[i];
                            cnt[4 * i] = v.x;
                            cnt[4 * i + 1] = v.y;
                            cnt[4 * i + 2] = v.z;
                            cnt[4 * i + 3] = v.w;
                        }
                        int cum = 0, nl = -1;
#pragma unroll
                        for (int t = NBIN; t >= 0; t--) {
                            cum += cnt[t];
                            nl = (nl < 0 && cum >= kL) ? t : nl;
                        }
                        if (nl > 0) {
                            const int old = atomicMax(&s_lvl[L], nl);
                            if (nl > old) {
                                const float b2 = s_qb[L].y * (1.0f - (float)nl * (BSPAN / (float)NBIN));
                                s_qrec[L].x = 0.5f * ((1.0f - c1) * qo.z - b2 * 1.000001f);
                                // published for the other blocks of this query block
                                atomicMax(&pcnt[(int64_t)(__float_as_int(qo.w) - q0) * PCS + 1], nl);
                            }
                        }
                    }
                }
#pragma unroll
                for (int a = 0; a < 4; a++)
#pragma unroll
                    for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
            }
            // both counters written on every path: a store to one of two
            // counters chosen by the branch would keep them in scratch memory,
            // and each scratch load waits for every load in flight
            const bool last = ck == nk - 1;
            cj += last ? 1 : 0;
            ck = last ? 0 : ck + 1;
        };
        uint4 buf[CNB][3];
#pragma unroll
        for (int b = 0; b < CNB; b++) ld(buf[b]);
        __builtin_amdgcn_sched_barrier(0);
        for (int s0 = 0; s0 < nsteps; s0 += CNB) {
#pragma unroll
            for (int b = 0; b < CNB; b++) {
                if (s0 + b < nsteps) step(buf[b][0], buf[b][1], buf[b][2]);
                __builtin_amdgcn_sched_barrier(0);
                ld(buf[b]);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    }
    // totals with one atomic per block: the blocks end together
    block_add_u64(nver, verified);
    if (ncand) block_add_u64(ncnd, ncand);
}

template <typename T>
void dfree_(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

template <typename T>
T *dalloc_(int64_t n)
{
    T *p = nullptr;
    if (hipMalloc((void **)&p, (size_t)std::max<int64_t>(n, 1) * sizeof(T)) != hipSuccess)
        fail(GTS_E_OOM, "cell scan arrays");
    return p;
}

// One cell per 256 rows (8 chunks), at most 4,096.  Fewer, larger cells put
// fewer chunks on cell boundaries (1M x 300 with 1,000 clusters: 36.4-36.8k
// live chunks at 2,000-3,000 cells, 38.1k at 3,906) until a cluster has
// fewer than about two cells, where the bounds loosen (60k x 300 with 120
// clusters: 133 cells verify 140M pairs, 234 cells under half of 180M), so
// the count stays on the small side of that edge.
int auto_cells(int64_t n)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>(4096, n / 256));
}

// nearest of the P pivot rows (tree positions cp) for every row: cell, distance
void assign_rows(gts_index *I, const int32_t *cp, int P, const int32_t *rows, int32_t *cell, float *dist)
{
    float *D = I->dense.ensure(ASSIGN_ROWS * P);
    for (int64_t r0 = 0; r0 < I->n; r0 += ASSIGN_ROWS) {
        const int64_t nr = std::min<int64_t>(ASSIGN_ROWS, I->n - r0);
        pivot_table(I, rows + r0, nr, cp, P, D);
        k_cell_argmin<<<(unsigned)((nr + 3) / 4), 256, 0, I->st>>>(D, nr, P, -1, cell + r0, dist + r0);
        GTS_HIP(hipGetLastError());
    }
}

}  // namespace

void gather_split_rows(gts_index *I, const int32_t *pos, int64_t n, int64_t npad, uint16_t *hi, uint16_t *lo,
                       float *n2)
{
    ensure_bf16_split(I);
    hipStream_t st = I->st;
    k_scan_split<<<grid_for(npad * (I->dk / 8), 256), 256, 0, st>>>(I->Xhi, I->Xlo, pos, n, npad, I->dk, hi, lo);
    k_gather_f32<<<grid_for(n, 256), 256, 0, st>>>(I->xn2, pos, n, n2);
    GTS_HIP(hipGetLastError());
}

bool cells_supported(const gts_index *I)
{
    return I->metric == GTS_METRIC_L2 && I->verify_form == GTS_VERIFY_BF16X3 && I->death_ptr == nullptr &&
           ((I->dp + 31) / 32) * 32 <= 512 && (I->n + DR - 1) / DR <= (int64_t)32 * 1024 * 32;
}

void cells_release(gts_index *I)
{
    CellIndex &c = I->cells;
    dfree_(c.cpos);
    dfree_(c.scan_pos);
    dfree_(c.pos_scan);
    dfree_(c.sid);
    dfree_(c.xn2);
    dfree_(c.rows);
    dfree_(c.eps);
    c.eps_max = 0.0f;
    dfree_(c.cell_rows);
    dfree_(c.sdist);
    dfree_(c.phi);
    dfree_(c.pn2);
    dfree_(c.cell_pack);
    c.built = false;
    c.ncell = 0;
}

void cells_build(gts_index *I)
{
    CellIndex &C = I->cells;
    if (C.built) return;
    if (I->metric != GTS_METRIC_L2 || I->n < 1) fail(GTS_E_ARG, "the cell scan needs L2 rows");
    const auto t_start = std::chrono::steady_clock::now();
    ensure_bf16_split(I);
    hipStream_t st = I->st;
    const int64_t n = I->n;
    const int P = (int)std::min<int64_t>(n, C.want > 0 ? C.want : auto_cells(n));
    // strided pivots first, then FAR_ROUNDS rounds of far rows (a quarter of
    // the pivots in all)
    const int P0 = std::max(1, P - P / 4);
    // pivot rows: a strided sample of tree positions
    std::vector<int32_t> hcp(P0);
    for (int i = 0; i < P0; i++) hcp[i] = (int32_t)((int64_t)i * n / P0);
    int32_t *rows = dalloc_<int32_t>(n), *cell = dalloc_<int32_t>(n), *cp = dalloc_<int32_t>(P);
    float *dist = dalloc_<float>(n);
    fill_iota_i32(rows, n, st);
    GTS_HIP(hipMemcpyAsync(cp, hcp.data(), P0 * 4, hipMemcpyHostToDevice, st));
    assign_rows(I, cp, P0, rows, cell, dist);
    // then rows far from every pivot so far (beyond twice the median
    // distance), evenly spaced, become pivots of cells of their own, in
    // rounds: a cluster that the strided sample missed sends all its rows
    // into one distant cell, and one round can spend its pivots on the far
    // tails of covered clusters instead (1M x 300, 2,500 cells: 79,772 live
    // chunks after one round, 37,088 after three)
    int Pn = P0;
    for (int round = 0; round < FAR_ROUNDS && P > Pn; round++) {
        const int target = round + 1 == FAR_ROUNDS ? P : Pn + (P - P0) / FAR_ROUNDS;
        std::vector<float> hd(n);
        GTS_HIP(hipMemcpyAsync(hd.data(), dist, n * 4, hipMemcpyDeviceToHost, st));
        GTS_HIP(hipStreamSynchronize(st));
        std::vector<float> tmp(hd);
        std::nth_element(tmp.begin(), tmp.begin() + n / 2, tmp.end());
        const float thr = 2.0f * tmp[n / 2];
        std::vector<int32_t> far;
        for (int64_t i = 0; i < n; i++)
            if (hd[i] > thr) far.push_back((int32_t)i);
        const int64_t add = std::min<int64_t>(target - Pn, (int64_t)far.size());
        for (int64_t i = 0; i < add; i++) hcp.push_back(far[(size_t)(i * (int64_t)far.size() / add)]);
        if ((int)hcp.size() == Pn) break;
        Pn = (int)hcp.size();
        GTS_HIP(hipMemcpyAsync(cp, hcp.data(), Pn * 4, hipMemcpyHostToDevice, st));
        assign_rows(I, cp, Pn, rows, cell, dist);
    }
    // cell numbering: connected components of the pivots' nearest-neighbour
    // graph get consecutive numbers (cells of one cluster are adjacent)
    std::vector<int32_t> rank(Pn);
    {
        std::vector<int32_t> nn(Pn, 0);
        if (Pn > 1) {
            float *D = I->dense.ensure((int64_t)Pn * Pn);
            int32_t *dnn = dalloc_<int32_t>(Pn);
            pivot_table(I, cp, Pn, cp, Pn, D);
            k_cell_argmin<<<(unsigned)((Pn + 3) / 4), 256, 0, st>>>(D, Pn, Pn, 0, dnn, nullptr);
            GTS_HIP(hipGetLastError());
            GTS_HIP(hipMemcpyAsync(nn.data(), dnn, Pn * 4, hipMemcpyDeviceToHost, st));
            GTS_HIP(hipStreamSynchronize(st));
            dfree_(dnn);
        }
        std::vector<int32_t> par(Pn);
        std::iota(par.begin(), par.end(), 0);
        auto find = [&](int32_t a) {
            while (par[a] != a) {
                par[a] = par[par[a]];
                a = par[a];
            }
            return a;
        };
        for (int32_t a = 0; a < Pn; a++) {
            const int32_t ra = find(a), rb = find(nn[a]);
            if (ra != rb) par[std::max(ra, rb)] = std::min(ra, rb);
        }
        std::vector<int32_t> order(Pn);
        std::iota(order.begin(), order.end(), 0);
        std::vector<int32_t> comp(Pn);
        for (int32_t a = 0; a < Pn; a++) comp[a] = find(a);
        std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return comp[a] < comp[b]; });
        for (int32_t r = 0; r < Pn; r++) rank[order[r]] = r;
        std::vector<int32_t> sorted_cp(Pn);
        for (int32_t a = 0; a < Pn; a++) sorted_cp[rank[a]] = hcp[a];
        hcp = sorted_cp;
    }
    // scan order: rows sorted by (cell number, distance to the pivot)
    int32_t *drank = dalloc_<int32_t>(Pn);
    GTS_HIP(hipMemcpyAsync(drank, rank.data(), Pn * 4, hipMemcpyHostToDevice, st));
    uint64_t *k0 = dalloc_<uint64_t>(n), *k1 = dalloc_<uint64_t>(n);
    int32_t *v0 = dalloc_<int32_t>(n), *v1 = dalloc_<int32_t>(n);
    k_scan_keys<<<grid_for(n, 256), 256, 0, st>>>(cell, dist, drank, n, k0, v0);
    GTS_HIP(hipGetLastError());
    int cbits = 0;
    while ((1 << cbits) < Pn) cbits++;
    uint64_t *ks = nullptr;
    int32_t *vs = nullptr;
    radix_sort_u64(k0, v0, k1, v1, n, 32 + cbits, nullptr, 0, st, &ks, &vs);
    const int64_t npad = (n + DR - 1) / DR * DR, nchunk = npad / DR;
    C.ncell = Pn;
    C.cpos = dalloc_<int32_t>(Pn);
    GTS_HIP(hipMemcpyAsync(C.cpos, hcp.data(), Pn * 4, hipMemcpyHostToDevice, st));
    C.scan_pos = dalloc_<int32_t>(n);
    GTS_HIP(hipMemcpyAsync(C.scan_pos, vs, n * 4, hipMemcpyDeviceToDevice, st));
    C.pos_scan = dalloc_<int32_t>(n);
    C.sid = dalloc_<int32_t>(n);
    C.xn2 = dalloc_<float>(n);
    int32_t *scell = cell;   // reused: cell of each scan row
    k_scan_rows<<<grid_for(n, 256), 256, 0, st>>>(ks, C.scan_pos, n, I->perm, I->xn2, C.pos_scan, C.sid, C.xn2,
                                                  scell, dist);
    GTS_HIP(hipGetLastError());
    // the hi and FP8 planes first apart, then interleaved by K step
    uint16_t *hi = dalloc_<uint16_t>(npad * I->dk);
    uint8_t *lo8 = dalloc_<uint8_t>(npad * I->dk);
    k_scan_split<<<grid_for(npad * (I->dk / 8), 256), 256, 0, st>>>(I->Xhi, I->Xlo, C.scan_pos, n, npad, I->dk,
                                                                    hi, nullptr);
    GTS_HIP(hipGetLastError());
    // FP8 residuals, scaled so that the largest possible one, 2^-8 max|x_i|
    // <= 2^-8 max|x|, maps to at most 240 (within both FP8 E4M3 variants)
    {
        const double xmax = std::sqrt((double)I->xn2_max);
        int shift = 0;
        if (xmax > 0.0) shift = (int)std::floor(std::log2(240.0 * 256.0 / xmax));
        C.lo_shift = std::max(-100, std::min(100, shift));
        C.eps = dalloc_<float>(npad);
        k_scan_lo8<<<grid_for(npad, 256), 256, 0, st>>>(I->X, I->dp, I->dim, I->dk, C.scan_pos, n, npad, C.lo_shift,
                                                        lo8, C.eps);
        GTS_HIP(hipGetLastError());
        C.rows = dalloc_<uint8_t>(npad * I->dk * 3);
        k_interleave_rows<<<grid_for(nchunk * (I->dk / 32) * 192, 256), 256, 0, st>>>(
            hi, lo8, nchunk, I->dk / 32, C.rows);
        GTS_HIP(hipGetLastError());
        std::vector<float> he(npad);
        GTS_HIP(hipMemcpyAsync(he.data(), C.eps, npad * 4, hipMemcpyDeviceToHost, st));
        GTS_HIP(hipStreamSynchronize(st));
        dfree_(hi);
        dfree_(lo8);
        C.eps_max = 0.0f;
        for (float v : he) C.eps_max = std::max(C.eps_max, v);
    }
    C.cell_rows = dalloc_<int2>(Pn);
    fill_i32((int32_t *)C.cell_rows, -1, 2 * (int64_t)Pn, st);
    k_cell_ranges<<<grid_for(n, 256), 256, 0, st>>>(scell, n, C.cell_rows);
    GTS_HIP(hipGetLastError());
    // pivot rows for the bounds kernel: BF16 hi copies in cell order, norms
    const int64_t ppad = ((int64_t)Pn + 31) / 32 * 32;
    C.phi = dalloc_<uint16_t>(ppad * I->dk);
    k_scan_split<<<grid_for(ppad * (I->dk / 8), 256), 256, 0, st>>>(I->Xhi, I->Xlo, C.cpos, Pn, ppad, I->dk, C.phi,
                                                                    nullptr);
    C.pn2 = dalloc_<float>(Pn);
    k_gather_f32<<<grid_for(Pn, 256), 256, 0, st>>>(I->xn2, C.cpos, Pn, C.pn2);
    C.cell_pack = dalloc_<float4>(Pn);
    k_cell_pack<<<grid_for(Pn, 256), 256, 0, st>>>(C.pn2, C.cell_rows, dist, Pn, I->em, C.cell_pack);
    GTS_HIP(hipGetLastError());
    GTS_HIP(hipStreamSynchronize(st));
    C.sdist = dist;   // distance of each scan row to its cell's pivot row
    for (void *p : {(void *)rows, (void *)cell, (void *)cp, (void *)drank, (void *)k0, (void *)k1, (void *)v0,
                    (void *)v1})
        (void)hipFree(p);
    C.built = true;
    const std::chrono::duration<double, std::milli> dt = std::chrono::steady_clock::now() - t_start;
    I->stats.ms_cells_build += dt.count();
}

void cells_sort_queries(gts_index *I, int32_t *qpos, int32_t *kq, float *bnd, int32_t *orig, int64_t nq)
{
    hipStream_t st = I->st;
    int32_t *perm = I->qperm.ensure(nq), *t = I->qtmp.ensure(3 * nq);
    float *tf = I->qtmp_f.ensure(nq);
    const int64_t nseg = (nq + QSORT - 1) / QSORT;
    int32_t *rank = I->qrank.ensure(nq);
    fill_i32(rank, 0, nq, st);
    const int64_t len = std::min<int64_t>(nq, QSORT);
    const int NI = (int)((len + 255) / 256), NJ = (int)((len + RTILE - 1) / RTILE);
    k_qrank<<<(unsigned)(nseg * NI * NJ), 256, 0, st>>>(qpos, nullptr, I->cells.pos_scan, nq, I->n, NI, NJ, rank);
    k_qplace<<<grid_for(nq, 256), 256, 0, st>>>(rank, nq, perm);
    k_qgather<<<grid_for(nq, 256), 256, 0, st>>>(perm, nq, qpos, kq, bnd, orig, t, tf);
    k_qscatter<<<grid_for(nq, 256), 256, 0, st>>>(nq, t, tf, qpos, kq, bnd, orig);
    GTS_HIP(hipGetLastError());
    I->qrank_zero = nullptr;   // ranks left in place: cleared before the next cells_prepare_batch
}

void cells_prepare_batch(gts_index *I, const int32_t *qids, const int32_t *ks, int kfix, int kcap, int32_t *argerr,
                         int64_t nq, int32_t *qpos, int32_t *kq, int32_t *orig, float *bnd, int32_t *ccount,
                         int32_t *scal8)
{
    hipStream_t st = I->st;
    int64_t nbits = 0;
    uint32_t *bits = cells_bitmap(I, nq, &nbits);
    int32_t *rank = I->qrank.ensure(nq);
    if (I->qrank_zero != rank) fail(GTS_E_ARG, "query ranks not cleared (cells_ranks_ready)");
    const int64_t nseg = (nq + QSORT - 1) / QSORT, len = std::min<int64_t>(nq, QSORT);
    const int NI = (int)((len + 255) / 256), NJ = (int)((len + RTILE - 1) / RTILE);
    k_qrank<<<(unsigned)(nseg * NI * NJ), 256, 0, st>>>(qids, I->inv, I->cells.pos_scan, nq, I->n, NI, NJ, rank);
    k_qplace_fields<<<grid_for(std::max(nq, nbits), 256), 256, 0, st>>>(rank, qids, I->inv, I->n, ks, kfix, kcap,
                                                                        argerr, nq, qpos, kq, orig, bnd, ccount,
                                                                        scal8, bits, nbits);
    GTS_HIP(hipGetLastError());
}

void cells_ranks_ready(gts_index *I, int64_t nq)
{
    int32_t *rank = I->qrank.ensure(nq);
    if (I->qrank_zero == rank && I->qrank_zero_cap == I->qrank.cap) return;
    GTS_HIP(hipMemsetAsync(rank, 0, I->qrank.cap * sizeof(int32_t), I->st));
    I->qrank_zero = rank;
    I->qrank_zero_cap = I->qrank.cap;
}

uint32_t *cells_bitmap(gts_index *I, int64_t nq, int64_t *nwords)
{
    const int64_t nchunk = (I->n + DR - 1) / DR, nw = (nchunk + 31) / 32;
    const int64_t sqb = (std::min(nq, QSEG) + QB - 1) / QB;
    *nwords = sqb * nw;
    return I->cell_bits.ensure(sqb * nw);
}

// the verification kernel's padded per-query counters for one segment of
// QSEG queries, zero whenever no pass is running (cleared when allocated,
// left clear by k_cell_counts); allocated outside graph capture by
// cells_reserve
int32_t *cell_counters(gts_index *I)
{
    const bool fresh = I->cell_pcnt.p == nullptr || I->cell_pcnt.cap < (size_t)QSEG * PCS;
    int32_t *p = I->cell_pcnt.ensure((size_t)QSEG * PCS);
    if (fresh) fill_i32(p, 0, (int64_t)QSEG * PCS, I->st);
    return p;
}

void cells_reserve(gts_index *I, int64_t nq, void (*mix)(void *, uint64_t), void *ctx)
{
    const CellIndex &C = I->cells;
    const int64_t nchunk = (I->n + DR - 1) / DR;
    const int64_t sq = std::min(nq, QSEG), sqb = (sq + QB - 1) / QB;
    const void *p[] = {I->cell_bits.ensure(sqb * ((nchunk + 31) / 32)), I->cell_list.ensure(sqb * nchunk),
                       I->cell_cnt.ensure(sqb), I->qperm.ensure(nq),
                       I->qrank.ensure(nq), I->qtmp.ensure(3 * nq), I->qtmp_f.ensure(nq),
                       stage_query_image(I, nullptr, nq), C.rows, C.cpos, C.pos_scan, cell_counters(I)};
    for (const void *v : p) mix(ctx, (uint64_t)(uintptr_t)v);
}

void cells_launch(gts_index *I, const Search &S, const VerifyArgs &va, hipEvent_t ev0, hipEvent_t ev1,
                  bool external, int ready, bool rerun)
{
    const bool staged = (ready & CELLS_STAGED) != 0;
    const CellIndex &C = I->cells;
    if (!C.built) fail(GTS_E_ARG, "the cell scan is not built");
    hipStream_t st = I->st;
    const int64_t nchunk = (I->n + DR - 1) / DR, nw = (nchunk + 31) / 32;
    const int P = C.ncell;
    const BfErr be = bf_err_model(I->dim, I->dk, 3);
    const uint16_t *qimg = stage_query_image(I, staged ? nullptr : S.qpos, S.nq);
    const size_t lds_img = (size_t)2 * QB * (I->dk + 8) * 2;
    // LDS beside the query image and the work offsets (16 KB left for the
    // static arrays): the row terms of up to cs chunks (12 B per row), cs a
    // multiple of the wave count
    const int64_t budget = (int64_t)160 * 1024 - 16384 - (int64_t)lds_img - (QSEG / QB + 1) * 4;
    const int cs_max = (int)std::min<int64_t>(CS, budget / (DR * 12) / CW * CW);
    if (cs_max < CW) fail(GTS_E_CAPACITY, "cell scan tile exceeds LDS");
    const size_t lds = lds_img + (size_t)cs_max * DR * 12 + (size_t)(QSEG / QB + 1) * 4;
    // the bounds kernel: the image's hi half, and its bitmap row beside it
    // when that fits
    const size_t lds_hi = (size_t)QB * (I->dk + 8) * 2;
    const bool lds_bits = lds_hi + (size_t)nw * 4 <= (size_t)150 * 1024;
    const size_t lds_bounds = lds_hi + (lds_bits ? (size_t)nw * 4 : 0);
    const BfErr be_hi = bf_err_model(I->dim, I->dk, 1);
    (void)hipFuncSetAttribute((const void *)k_cell_bounds, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds_bounds);
    auto verify_kernel = rerun ? k_verify_cells_bf16<true> : k_verify_cells_bf16<false>;
    (void)hipFuncSetAttribute((const void *)verify_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    int ncu = 0;
    GTS_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, I->device));
    const unsigned grid = (unsigned)std::max(8, ncu / 8 * 8);
    int32_t *pcnt = cell_counters(I);
    auto rec = [&](hipEvent_t e) {
        if (e) GTS_HIP(external ? hipEventRecordWithFlags(e, st, hipEventRecordExternal) : hipEventRecord(e, st));
    };
    for (int64_t q0 = 0; q0 < S.nq; q0 += QSEG) {
        const int64_t sq = std::min(QSEG, S.nq - q0), sqb = (sq + QB - 1) / QB;
        // few query blocks (reruns): smaller bound tiles and work items, so
        // that the pass still spreads over the CUs
        const int bpass = sqb >= 32 ? BPASS : 1;
        const int cs = sqb >= 32 ? cs_max : CW;
        const int64_t npb = (P + BPT * bpass - 1) / (BPT * bpass);
        uint32_t *bits = I->cell_bits.ensure(sqb * nw);
        int32_t *list = I->cell_list.ensure(sqb * nchunk), *cnt = I->cell_cnt.ensure(sqb);
        if (!(q0 == 0 && (ready & CELLS_CLEARED))) fill_i32((int32_t *)bits, 0, sqb * nw, st);
        k_cell_bounds<<<(unsigned)(sqb * npb), BW * 64, lds_bounds, st>>>(
            qimg, I->dk, S.qpos, I->xn2, S.bnd, S.nq, q0, C.phi, C.cell_pack, P, C.cell_rows, C.sdist, nw, bpass,
            lds_bits, be_hi, I->em, bits);
        k_cell_compact<<<(unsigned)sqb, 256, 0, st>>>(bits, nw, nchunk, list, cnt);
        GTS_HIP(hipGetLastError());
        if (q0 == 0) rec(ev0);
        verify_kernel<<<grid, CW * 64, lds, st>>>(
            C.rows, ldexpf(1.0f, -C.lo_shift), C.eps, C.xn2, C.sid, I->n, I->dim, I->dk, list, nchunk, cnt, sqb, cs,
            q0, S.nq, qimg, S.qpos, I->xn2, S.bnd, S.kq, be, va.cand, pcnt, va.cap,
            (unsigned long long *)va.verified, (unsigned long long *)va.live, (unsigned long long *)va.ncand);
        GTS_HIP(hipGetLastError());
        if (q0 + QSEG >= S.nq) rec(ev1);   // the verification launches end here
        k_cell_counts<<<grid_for(sq, 256), 256, 0, st>>>(pcnt, sq, q0, va.ccount);
        GTS_HIP(hipGetLastError());
    }
}

}  // namespace gts
