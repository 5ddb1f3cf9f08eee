// Learned kNN radius (no counterpart in the reference).
//
// A 16 -> 64 -> 64 -> 1 network with exact GELU, BF16 weights and FP32
// accumulation maps a query's 16 smallest distances to the pivots of one tree
// level (log-scaled, standardised) to log of its k-th NN distance.  It is
// trained on the GPU by the kernels below (forward, backward, per-parameter
// gradient reduction, Adam on FP32 master weights) with labels from the exact
// plain kNN.  At query time the predicted radius, times a safety factor fitted
// on the training residuals, becomes the bound of one range pass; the top-k
// selection reports any query whose k-th exact distance is not within the
// radius, and the driver reruns those with a wider one, so results stay exact.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <random>

#include "gts_cells.h"
#include "gts_mfma.h"
#include "gts_api_internal.h"
#include "gts_dist.h"
#include "gts_search.h"
#include "gts_verify2.h"
#include "gts_wave.h"

namespace gts {

namespace {

constexpr int DIN = 16, DH = 64;
constexpr int NW = DH * DIN + DH + DH * DH + DH + DH + 1;  // 5313 parameters
constexpr int OW1 = 0, OB1 = OW1 + DH * DIN, OW2 = OB1 + DH, OB2 = OW2 + DH * DH,
              OW3 = OB2 + DH, OB3 = OW3 + DH;
constexpr int FEAT_MAXP = 4096;  // most feature pivots (bounds the query x pivot table)
// Feature pivots: the shallowest level with at least FEAT_MINP pivots.  On the
// 1M x 300 clustered rows the 100 pivots of level 2 predict as well as the
// 1,000 of level 3 (6 instead of 10 reruns per 10,000 queries) at a tenth of
// the distance table: features and network 53 -> 31 us per batch.
constexpr int FEAT_MINP = 64;
// row pitch of the transposed weight copies in LDS: DH + 1 words, so that the
// transposing stores (consecutive threads: consecutive rows) and the reads
// (consecutive lanes: consecutive columns) both hit distinct banks
constexpr int WP = DH + 1;

unsigned grid_for(int64_t n, int block)
{
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (unsigned)g;
}

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float x)
{
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * expf(-0.5f * x * x) * 0.3989422804014327f;
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ uint16_t f2bf(float v)
{
    uint32_t u = __float_as_uint(v);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Features: the 16 smallest distances to the level pivots, log-scaled and
// standardised.  The distances come from the tiled query x pivot table
// (pivot_table); one wave per query keeps the 16 smallest with the wave-level
// top-64 merge.
__global__ void __launch_bounds__(256) k_feat_select(const float *__restrict__ D, int64_t nq, int np,
                                                     const float *mu, const float *sd, float *feat)
{
    const int lane = threadIdx.x & 63;
    const int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= nq) return;   // the whole wave
    const float *dq = D + q * (int64_t)np;
    auto key = [&](int p) { return p < np ? ((uint64_t)__float_as_uint(dq[p]) << 32) | (uint32_t)p : ~0ull; };
    // The 16th smallest of the 64 lane minima bounds the 16th smallest key
    // (16 distinct keys lie at or below it); the keys up to that bound are
    // compacted into one 64-key slice and sorted.  More than 64 of them (rare)
    // fall back to a running top-64 over the row.
    __shared__ uint64_t s_sel[4][64];
    const int wv = threadIdx.x >> 6;
    uint64_t mn = ~0ull;
    for (int b0 = 0; b0 < np; b0 += 256) {
        uint64_t v[4];   // four slices in flight
#pragma unroll
        for (int u = 0; u < 4; u++) v[u] = key(b0 + 64 * u + lane);
#pragma unroll
        for (int u = 0; u < 4; u++) mn = min(mn, v[u]);
    }
    const uint64_t T = lane_key64(wave_sort64(mn, lane), DIN - 1);
    int nsel = 0;
    for (int b0 = 0; b0 < np; b0 += 64) {
        const uint64_t v = key(b0 + lane);
        const bool in = v <= T;
        const uint64_t m = __builtin_amdgcn_ballot_w64(in);
        const int pos = nsel + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
        if (in && pos < 64) s_sel[wv][pos] = v;
        nsel += __popcll(m);
    }
    uint64_t best;
    if (nsel <= 64) {
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        best = wave_sort64(lane < nsel ? s_sel[wv][lane] : ~0ull, lane);
    } else {
        best = ~0ull;
        for (int b0 = 0; b0 < np; b0 += 64) {
            const uint64_t v = key(b0 + lane);
            if (!__builtin_amdgcn_ballot_w64(v < lane_key64(best, DIN - 1))) continue;
            best = wave_merge64(best, v, lane);
        }
    }
    // fewer than 16 pivots: repeat the largest (every lane reads, then lanes < 16 write)
    const uint64_t mine = lane_read64(best, min(lane, np - 1));
    if (lane < DIN) {
        const float v = logf(__uint_as_float((uint32_t)(mine >> 32)) + 1e-6f);
        feat[q * DIN + lane] = mu ? (v - mu[lane]) / sd[lane] : v;
    }
}

// The network's BF16 weights into LDS as FP32, the two matrices transposed
// (pitch WP): the matrices read as 16-byte groups, every load of a thread in
// flight before its LDS stores (a load-then-store loop waits one latency per
// element).  Blocks of 128 threads or more.
__device__ __forceinline__ void stage_net(const uint16_t *__restrict__ wb, float *W1T, float *B1, float *W2T,
                                          float *B2, float *W3, float *b3)
{
    static_assert(OW1 % 8 == 0 && OW2 % 8 == 0 && (DH * DH) % 8 == 0 && (DH * DIN) % 8 == 0, "16-byte groups");
    constexpr int G2 = DH * DH / 8, G1 = DH * DIN / 8;   // 512, 128
    const int t = threadIdx.x, nt = blockDim.x;
    uint4 a[2], c = make_uint4(0, 0, 0, 0);
#pragma unroll
    for (int u = 0; u < 2; u++) a[u] = t + u * nt < G2 ? ((const uint4 *)(wb + OW2))[t + u * nt] : make_uint4(0, 0, 0, 0);
    if (t < G1) c = ((const uint4 *)(wb + OW1))[t];
    float v1 = 0.f, v2 = 0.f, v3 = 0.f, v4 = 0.f;
    if (t < DH) {
        v1 = bf2f(wb[OB1 + t]);
        v2 = bf2f(wb[OB2 + t]);
        v3 = bf2f(wb[OW3 + t]);
    }
    if (t == 0) v4 = bf2f(wb[OB3]);
    auto put = [&](const uint4 &g, int i0, float *WT, int din) {
        const uint32_t w[4] = {g.x, g.y, g.z, g.w};
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int i = i0 + e, u = i / din, j = i % din;
            WT[j * WP + u] = __uint_as_float((e & 1) ? (w[e >> 1] & 0xFFFF0000u) : (w[e >> 1] << 16));
        }
    };
#pragma unroll
    for (int u = 0; u < 2; u++)
        if (t + u * nt < G2) put(a[u], (t + u * nt) * 8, W2T, DH);
    if (t < G1) put(c, t * 8, W1T, DIN);
    if (t < DH) {
        B1[t] = v1;
        B2[t] = v2;
        W3[t] = v3;
    }
    if (t == 0) *b3 = v4;
    for (int i = 2 * nt + t; i < G2; i += nt) put(((const uint4 *)(wb + OW2))[i], i * 8, W2T, DH);   // blocks under 256
}

// Forward pass, one thread per sample; stores activations for backward.
__global__ void __launch_bounds__(256) k_mlp_fwd(const float *X, int64_t nb, const uint16_t *wb,
                                                 float *z1, float *a1, float *z2, float *a2,
                                                 float *y)
{
    __shared__ float W[NW];
    for (int i = threadIdx.x; i < NW; i += blockDim.x) W[i] = bf2f(wb[i]);
    __syncthreads();
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nb) return;
    float x[DIN], h1[DH];
#pragma unroll
    for (int j = 0; j < DIN; j++) x[j] = X[s * DIN + j];
#pragma unroll
    for (int i = 0; i < DH; i++) {
        float z = W[OB1 + i];
#pragma unroll
        for (int j = 0; j < DIN; j++) z = fmaf(W[OW1 + i * DIN + j], x[j], z);
        h1[i] = gelu(z);
        if (z1) {
            z1[s * DH + i] = z;
            a1[s * DH + i] = h1[i];
        }
    }
    float out = W[OB3];
    for (int i = 0; i < DH; i++) {
        float z = W[OB2 + i];
#pragma unroll
        for (int j = 0; j < DH; j++) z = fmaf(W[OW2 + i * DH + j], h1[j], z);
        float a = gelu(z);
        if (z2) {
            z2[s * DH + i] = z;
            a2[s * DH + i] = a;
        }
        out = fmaf(W[OW3 + i], a, out);
    }
    y[s] = out;
}

// Output error and hidden deltas of loss = mean (y - t)^2 / 2.
__global__ void __launch_bounds__(256) k_mlp_bwd(const float *t, int64_t nb, const uint16_t *wb,
                                                 const float *z1, const float *z2, const float *y,
                                                 float *e, float *d1, float *d2, float *loss)
{
    __shared__ float W[NW];
    __shared__ float red[256];
    for (int i = threadIdx.x; i < NW; i += blockDim.x) W[i] = bf2f(wb[i]);
    __syncthreads();
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (s < nb) {
        float err = y[s] - t[s];
        l = 0.5f * err * err;
        float es = err / (float)nb;
        e[s] = es;
        float dd2[DH];
#pragma unroll
        for (int i = 0; i < DH; i++) {
            dd2[i] = es * W[OW3 + i] * gelu_d(z2[s * DH + i]);
            d2[s * DH + i] = dd2[i];
        }
        for (int j = 0; j < DH; j++) {
            float acc = 0.0f;
#pragma unroll
            for (int i = 0; i < DH; i++) acc = fmaf(dd2[i], W[OW2 + i * DH + j], acc);
            d1[s * DH + j] = acc * gelu_d(z1[s * DH + j]);
        }
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0] / (float)nb);
}

// Gradient of every parameter: one thread per parameter, summed over samples.
__global__ void k_mlp_grad(const float *X, int64_t nb, const float *a1, const float *a2,
                           const float *e, const float *d1, const float *d2, float *g)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NW) return;
    float acc = 0.0f;
    if (p < OB1) {
        int i = p / DIN, j = p % DIN;
        for (int64_t s = 0; s < nb; s++) acc = fmaf(d1[s * DH + i], X[s * DIN + j], acc);
    } else if (p < OW2) {
        int i = p - OB1;
        for (int64_t s = 0; s < nb; s++) acc += d1[s * DH + i];
    } else if (p < OB2) {
        int i = (p - OW2) / DH, j = (p - OW2) % DH;
        for (int64_t s = 0; s < nb; s++) acc = fmaf(d2[s * DH + i], a1[s * DH + j], acc);
    } else if (p < OW3) {
        int i = p - OB2;
        for (int64_t s = 0; s < nb; s++) acc += d2[s * DH + i];
    } else if (p < OB3) {
        int j = p - OW3;
        for (int64_t s = 0; s < nb; s++) acc = fmaf(e[s], a2[s * DH + j], acc);
    } else {
        for (int64_t s = 0; s < nb; s++) acc += e[s];
    }
    g[p] = acc;
}

// Adam on the FP32 master weights; refreshes the BF16 copy.
__global__ void k_adam(float *w, const float *g, float *m, float *v, uint16_t *wb, float lr,
                       float c1, float c2)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NW) return;
    const float b1 = 0.9f, b2 = 0.999f;
    float gi = g[p];
    float mi = b1 * m[p] + (1.0f - b1) * gi;
    float vi = b2 * v[p] + (1.0f - b2) * gi * gi;
    m[p] = mi;
    v[p] = vi;
    float wi = w[p] - lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
    w[p] = wi;
    wb[p] = f2bf(wi);
}

__global__ void k_to_bf16(const float *w, uint16_t *wb, int n)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) wb[p] = f2bf(w[p]);
}

// radius = exp(y * y_sd + y_mu) * scale for queries with k <= kmax, else +inf
// Radius network inference: one wave per query, lane i = hidden unit i of
// both layers, the same operation order per unit as k_mlp_fwd (the output
// sum is a wave reduction).  Weights (BF16) as FP32 in LDS, W1 and W2
// transposed so that lanes read consecutive words.  With bnd: the radius
// exp(y y_sd + y_mu) scale (+inf for k > kmax or a non-finite value).
__global__ void __launch_bounds__(256) k_mlp_infer(const float *__restrict__ X, int64_t nb,
                                                   const uint16_t *__restrict__ wb, const int32_t *kq, int kmax,
                                                   float y_mu, float y_sd, float scale, float *y_out, float *bnd)
{
    __shared__ float W1T[DIN * WP], B1[DH], W2T[DH * WP], B2[DH], W3[DH];
    __shared__ float H1[4][DH];
    __shared__ float b3;
    stage_net(wb, W1T, B1, W2T, B2, W3, &b3);
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int64_t s = (int64_t)blockIdx.x * 4 + wv; s < nb; s += (int64_t)gridDim.x * 4) {
        const float xl = lane < DIN ? X[s * DIN + lane] : 0.0f;
        float z = B1[lane];
#pragma unroll
        for (int j = 0; j < DIN; j++) z = fmaf(W1T[j * WP + lane], __shfl(xl, j, 64), z);
        H1[wv][lane] = gelu(z);
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        float z2 = B2[lane];
#pragma unroll 16
        for (int j = 0; j < DH; j++) z2 = fmaf(W2T[j * WP + lane], H1[wv][j], z2);
        float o = W3[lane] * gelu(z2);
        for (int off = 32; off > 0; off >>= 1) o += __shfl_xor(o, off, 64);
        const float y = o + b3;
        if (lane == 0) {
            if (y_out) y_out[s] = y;
            if (bnd) {
                const float r = expf(y * y_sd + y_mu) * scale;
                bnd[s] = (kq[s] <= kmax && isfinite(r)) ? r : INFINITY;
            }
        }
        __builtin_amdgcn_wave_barrier();   // H1 reused by the wave's next query
    }
}

__global__ void k_radius(const float *y, const int32_t *kq, int64_t nq, int kmax, float y_mu,
                         float y_sd, float scale, float *bnd)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nq;
         i += (int64_t)gridDim.x * blockDim.x) {
        float r = expf(y[i] * y_sd + y_mu) * scale;
        bnd[i] = (kq[i] <= kmax && isfinite(r)) ? r : INFINITY;
    }
}

// Radius of each query in one kernel: its features as k_feat_select forms
// them (the table row held in registers, np <= 64 * FR) and the network's
// forward pass as k_mlp_infer computes it, so the radius is the same float.
// Blocks are persistent: the weights are staged once per block, each wave
// then takes queries s, s + 4 gridDim, ...
template <int FR>
__global__ void __launch_bounds__(256) k_feat_mlp(const float *__restrict__ D, int64_t nq, int np,
                                                  const float *mu, const float *sd,
                                                  const uint16_t *__restrict__ wb, const int32_t *kq, int kmax,
                                                  float y_mu, float y_sd, float scale, float *bnd)
{
    __shared__ float W1T[DIN * WP], B1[DH], W2T[DH * WP], B2[DH], W3[DH];
    __shared__ float H1[4][DH];
    __shared__ uint32_t s_sel[4][64];
    __shared__ float b3;
    stage_net(wb, W1T, B1, W2T, B2, W3, &b3);
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int ns = (np + 63) / 64;   // slices of the row, <= FR
    for (int64_t s = (int64_t)blockIdx.x * 4 + wv; s < nq; s += (int64_t)gridDim.x * 4) {
        const float *dq = D + s * (int64_t)np;
        // the distances' bits as 32-bit keys (non-negative floats order as
        // their bits): the 16 smallest values, in order, are the features
        // k_feat_select forms from (value, pivot) keys
        uint32_t key[FR];
#pragma unroll
        for (int u = 0; u < FR; u++) {
            const int p = u * 64 + lane;
            key[u] = (u < ns && p < np) ? __float_as_uint(dq[p]) : ~0u;
        }
        uint32_t mn = ~0u;
#pragma unroll
        for (int u = 0; u < FR; u++) mn = min(mn, key[u]);
        const uint32_t T = (uint32_t)__shfl((int)wave_sort32(mn, lane), DIN - 1, 64);
        int nsel = 0;
#pragma unroll
        for (int u = 0; u < FR; u++) {
            if (u >= ns) break;
            const bool in = key[u] <= T;
            const uint64_t m = __builtin_amdgcn_ballot_w64(in);
            const int pos = nsel + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
            if (in && pos < 64) s_sel[wv][pos] = key[u];
            nsel += __popcll(m);
        }
        uint32_t best;
        if (nsel <= 64) {
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            best = wave_sort32(lane < nsel ? s_sel[wv][lane] : ~0u, lane);
        } else {
            best = ~0u;
#pragma unroll
            for (int u = 0; u < FR; u++) {
                if (u >= ns) break;
                if (!__builtin_amdgcn_ballot_w64(key[u] < (uint32_t)__shfl((int)best, DIN - 1, 64))) continue;
                best = wave_merge32(best, key[u], lane);
            }
        }
        const uint32_t mine = (uint32_t)__shfl((int)best, min(lane, np - 1), 64);
        float xl = 0.0f;
        if (lane < DIN) {
            const float v = logf(__uint_as_float(mine) + 1e-6f);
            xl = mu ? (v - mu[lane]) / sd[lane] : v;
        }
        float z = B1[lane];
#pragma unroll
        for (int j = 0; j < DIN; j++) z = fmaf(W1T[j * WP + lane], __shfl(xl, j, 64), z);
        H1[wv][lane] = gelu(z);
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        float z2 = B2[lane];
#pragma unroll 16
        for (int j = 0; j < DH; j++) z2 = fmaf(W2T[j * WP + lane], H1[wv][j], z2);
        float o = W3[lane] * gelu(z2);
        for (int off = 32; off > 0; off >>= 1) o += __shfl_xor(o, off, 64);
        const float y = o + b3;
        if (lane == 0) {
            const float r = expf(y * y_sd + y_mu) * scale;
            bnd[s] = (kq[s] <= kmax && isfinite(r)) ? r : INFINITY;
        }
        __builtin_amdgcn_wave_barrier();   // H1 and s_sel reused by the wave's next query
    }
}

struct Scratch {
    float *p = nullptr;
    size_t n = 0;
    float *get(size_t want)
    {
        if (want > n) {
            if (p) (void)hipFree(p);
            p = nullptr;
            if (hipMalloc((void **)&p, want * 4) != hipSuccess) fail(GTS_E_OOM, "network scratch");
            n = want;
        }
        return p;
    }
    ~Scratch()
    {
        if (p) (void)hipFree(p);
    }
};

constexpr int64_t FEAT_CH = 16384;   // queries per table (at most 16384 x 4096 x 4 B = 256 MB)

void features(gts_index *I, const int32_t *qpos, int64_t nq, const float *mu, const float *sd,
              float *feat)
{
    const LearnedModel &m = I->model;
    constexpr int64_t CH = FEAT_CH;
    for (int64_t q0 = 0; q0 < nq; q0 += CH) {
        const int64_t n = std::min(CH, nq - q0);
        float *D = I->ftab.ensure(n * m.npiv);
        if (m.phi) pivot_table_bf16(I, qpos + q0, n, m.phi, m.pn2, m.npiv, D);
        else pivot_table(I, qpos + q0, n, m.piv, m.npiv, D);
        k_feat_select<<<(unsigned)((n + 3) / 4), 256, 0, I->st>>>(D, n, m.npiv, mu, sd, feat + q0 * DIN);
        GTS_HIP(hipGetLastError());
    }
}

// The allocation-free loss and gradient (gts_mlp_loss_grad_noalloc) keeps the
// activations, deltas and BF16 weights of up to NOALLOC_NB samples in
// `scratch`, carved in 256-byte aligned pieces.
constexpr int64_t NOALLOC_NB = GTS_MLP_NOALLOC_MAX_NB;
constexpr int64_t al64(int64_t n) { return (n + 63) / 64 * 64; }
constexpr int64_t NOALLOC_FLOATS = 6 * al64(NOALLOC_NB * DH) + 2 * al64(NOALLOC_NB) + al64((NW + 1) / 2);

}  // namespace

// One copy per GPU the library runs on, held until the process ends; every
// call on a GPU shares it (calls that overlap on one GPU race).  Host code
// uses only the per-device addresses from hipGetSymbolAddress.
static __device__ __align__(256) float scratch[NOALLOC_FLOATS];

namespace {

constexpr int NOALLOC_DEVICES = 64;
std::mutex g_noalloc_mu;
float *g_noalloc_ptr[NOALLOC_DEVICES];

// the device address of this GPU's `scratch` (the GPU current), looked up
// once per device
float *noalloc_scratch(int device)
{
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (device < 0 || device >= NOALLOC_DEVICES) fail(GTS_E_DEVICE, "device ordinal out of range");
    if (!g_noalloc_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            fail(GTS_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_noalloc_ptr[device] = (float *)p;
    }
    return g_noalloc_ptr[device];
}

}  // namespace

int mlp_num_params() { return NW; }

void mlp_loss_grad_noalloc(gts_index *I, const float *X, const float *t, int64_t nb, const float *w, float *g,
                           float *loss)
{
    if (nb < 1 || nb > NOALLOC_NB)
        fail(GTS_E_ARG, "nb must be in [1, " + std::to_string(NOALLOC_NB) + "] (GTS_MLP_NOALLOC_MAX_NB)");
    hipStream_t st = I->st;
    float *base = noalloc_scratch(I->device);
    int64_t off = 0;
    auto take = [&](int64_t n) {
        float *p = base + off;
        off += al64(n);
        return p;
    };
    float *z1 = take(nb * DH), *a1 = take(nb * DH), *z2 = take(nb * DH), *a2 = take(nb * DH);
    float *d1 = take(nb * DH), *d2 = take(nb * DH), *y = take(nb), *e = take(nb);
    uint16_t *wb = (uint16_t *)take((NW + 1) / 2);
    if (off > NOALLOC_FLOATS) fail(GTS_E_ARG, "scratch layout exceeds its array");   // not reached: nb checked
    // every piece is written before it is read; the loss accumulates atomically
    // into the caller's word, cleared first (a kernel, no memset)
    fill_f32(loss, 0.0f, 1, st);
    k_to_bf16<<<grid_for(NW, 256), 256, 0, st>>>(w, wb, NW);
    k_mlp_fwd<<<grid_for(nb, 256), 256, 0, st>>>(X, nb, wb, z1, a1, z2, a2, y);
    k_mlp_bwd<<<grid_for(nb, 256), 256, 0, st>>>(t, nb, wb, z1, z2, y, e, d1, d2, loss);
    k_mlp_grad<<<grid_for(NW, 256), 256, 0, st>>>(X, nb, a1, a2, e, d1, d2, g);
    GTS_HIP(hipGetLastError());
    GTS_HIP(hipStreamSynchronize(st));
}

// loss and gradients for packed FP32 weights (BF16-rounded in the forward pass)
float mlp_loss_grad(gts_index *I, const float *hX, const float *ht, int64_t nb, const float *hw,
                    float *hg)
{
    hipStream_t st = I->st;
    Scratch sc;
    size_t off = 0;
    float *base = sc.get((size_t)nb * (DIN + 3 + 6 * DH) + 2 * NW + 1024);
    auto take = [&](size_t n) { float *p = base + off; off += (n + 63) / 64 * 64; return p; };
    float *X = take(nb * DIN), *t = take(nb), *z1 = take(nb * DH), *a1 = take(nb * DH);
    float *z2 = take(nb * DH), *a2 = take(nb * DH), *y = take(nb), *e = take(nb);
    float *d1 = take(nb * DH), *d2 = take(nb * DH), *w = take(NW), *g = take(NW), *loss = take(1);
    uint16_t *wb = nullptr;
    GTS_HIP(hipMalloc((void **)&wb, NW * 2));
    GTS_HIP(hipMemcpyAsync(X, hX, nb * DIN * 4, hipMemcpyHostToDevice, st));
    GTS_HIP(hipMemcpyAsync(t, ht, nb * 4, hipMemcpyHostToDevice, st));
    GTS_HIP(hipMemcpyAsync(w, hw, NW * 4, hipMemcpyHostToDevice, st));
    GTS_HIP(hipMemsetAsync(loss, 0, 4, st));
    k_to_bf16<<<grid_for(NW, 256), 256, 0, st>>>(w, wb, NW);
    k_mlp_fwd<<<grid_for(nb, 256), 256, 0, st>>>(X, nb, wb, z1, a1, z2, a2, y);
    k_mlp_bwd<<<grid_for(nb, 256), 256, 0, st>>>(t, nb, wb, z1, z2, y, e, d1, d2, loss);
    k_mlp_grad<<<grid_for(NW, 256), 256, 0, st>>>(X, nb, a1, a2, e, d1, d2, g);
    GTS_HIP(hipGetLastError());
    float l = 0;
    GTS_HIP(hipMemcpyAsync(hg, g, NW * 4, hipMemcpyDeviceToHost, st));
    GTS_HIP(hipMemcpyAsync(&l, loss, 4, hipMemcpyDeviceToHost, st));
    GTS_HIP(hipStreamSynchronize(st));
    (void)hipFree(wb);
    return l;
}

float learned_train(gts_index *I, const int32_t *train_qids, int64_t ntrain, int kmax, int epochs,
                    float lr, uint64_t seed)
{
    if (I->metric == GTS_METRIC_EDIT) fail(GTS_E_ARG, "the radius network needs a vector metric");
    if (kmax > I->n) fail(GTS_E_ARG, "kmax above the object count: every label would be infinite");
    if (kmax > topk_cap_limit()) fail(GTS_E_ARG, "kmax above GTS_MAX_K");
    hipStream_t st = I->st;
    LearnedModel &m = I->model;
    m.trained = false;
    I->plan_memo.clear();
    graph_reset(I);
    // feature pivots: the shallowest level with at least FEAT_MINP split
    // nodes (at most FEAT_MAXP)
    const std::vector<int32_t> best = level_pivots(I, FEAT_MAXP, &m.level, FEAT_MINP);
    if ((int)best.size() < DIN) fail(GTS_E_ARG, "tree too small for the radius network (fewer than 16 pivots per level)");
    if (m.piv) (void)hipFree(m.piv);
    GTS_HIP(hipMalloc((void **)&m.piv, best.size() * 4));
    GTS_HIP(hipMemcpyAsync(m.piv, best.data(), best.size() * 4, hipMemcpyHostToDevice, st));
    m.npiv = (int)best.size();
    for (void *p : {(void *)m.phi, (void *)m.pn2})
        if (p) (void)hipFree(p);
    m.phi = nullptr;
    m.pn2 = nullptr;
    if (I->metric == GTS_METRIC_L2 && ((I->dp + 31) / 32) * 32 <= 512) {
        // L2 features from MFMA distance tables: the pivots' split copies
        ensure_bf16_split(I);
        const int64_t ppad = ((int64_t)m.npiv + 31) / 32 * 32;
        GTS_HIP(hipMalloc((void **)&m.phi, ppad * I->dk * 2));
        GTS_HIP(hipMalloc((void **)&m.pn2, m.npiv * 4));
        gather_split_rows(I, m.piv, m.npiv, ppad, m.phi, nullptr, m.pn2);
    }

    // labels: exact k-th NN distances from the plain kNN (buffers released on
    // every path, errors included)
    Scratch sdq, sdi, sdd, sdk;
    int32_t *dq = (int32_t *)sdq.get(ntrain), *di = (int32_t *)sdi.get(ntrain * (size_t)kmax);
    float *dd = sdd.get(ntrain * (size_t)kmax), *dk = sdk.get(ntrain);
    GTS_HIP(hipMemcpyAsync(dq, train_qids, ntrain * 4, hipMemcpyHostToDevice, st));
    // (the plain path: a forced learned-only plan does not apply to it)
    const int saved_form = I->verify_form, saved_plan = I->knn_plan;
    I->knn_plan = GTS_PLAN_AUTO;
    try {
        knn_run(I, dq, ntrain, nullptr, kmax, kmax, di, dd, dk, 0, GTS_VERIFY_FP32);
    } catch (...) {
        I->verify_form = saved_form;
        I->knn_plan = saved_plan;
        throw;
    }
    I->verify_form = saved_form;
    I->knn_plan = saved_plan;
    std::vector<float> hk(ntrain);
    GTS_HIP(hipMemcpyAsync(hk.data(), dk, ntrain * 4, hipMemcpyDeviceToHost, st));

    // raw features of the training queries (positions = inv[qid])
    Scratch sqp, sfe;
    int32_t *qp = (int32_t *)sqp.get(ntrain);
    std::vector<int32_t> hinv(I->n);
    GTS_HIP(hipMemcpyAsync(hinv.data(), I->inv, I->n * 4, hipMemcpyDeviceToHost, st));
    GTS_HIP(hipStreamSynchronize(st));
    std::vector<int32_t> hqp(ntrain);
    for (int64_t i = 0; i < ntrain; i++) hqp[i] = hinv[train_qids[i]];
    GTS_HIP(hipMemcpyAsync(qp, hqp.data(), ntrain * 4, hipMemcpyHostToDevice, st));
    float *feat = sfe.get(ntrain * DIN);
    features(I, qp, ntrain, nullptr, nullptr, feat);
    std::vector<float> hf(ntrain * DIN);
    GTS_HIP(hipMemcpyAsync(hf.data(), feat, hf.size() * 4, hipMemcpyDeviceToHost, st));
    GTS_HIP(hipStreamSynchronize(st));
    for (int64_t i = 0; i < ntrain; i++)
        if (!std::isfinite(hk[i]))
            fail(GTS_E_ARG, "a training query's k-th NN distance is not finite (fewer than kmax live objects)");

    // standardisation constants (16 feature means / deviations, label mean / deviation)
    std::vector<float> mu(DIN, 0.f), sd(DIN, 0.f);
    std::vector<float> ty(ntrain);
    for (int j = 0; j < DIN; j++) {
        double s = 0, s2 = 0;
        for (int64_t i = 0; i < ntrain; i++) s += hf[i * DIN + j];
        s /= ntrain;
        for (int64_t i = 0; i < ntrain; i++) s2 += (hf[i * DIN + j] - s) * (hf[i * DIN + j] - s);
        mu[j] = (float)s;
        sd[j] = (float)std::max(std::sqrt(s2 / ntrain), 1e-6);
    }
    double ym = 0, ys = 0;
    for (int64_t i = 0; i < ntrain; i++) {
        ty[i] = std::log(std::max(hk[i], 1e-12f));
        ym += ty[i];
    }
    ym /= ntrain;
    for (int64_t i = 0; i < ntrain; i++) ys += (ty[i] - ym) * (ty[i] - ym);
    ys = std::max(std::sqrt(ys / ntrain), 1e-6);
    for (int64_t i = 0; i < ntrain; i++) {
        ty[i] = (float)((ty[i] - ym) / ys);
        for (int j = 0; j < DIN; j++) hf[i * DIN + j] = (hf[i * DIN + j] - mu[j]) / sd[j];
    }
    m.y_mu = (float)ym;
    m.y_sd = (float)ys;
    if (m.mu) (void)hipFree(m.mu);
    GTS_HIP(hipMalloc((void **)&m.mu, 2 * DIN * 4));
    m.sd = m.mu + DIN;
    GTS_HIP(hipMemcpyAsync(m.mu, mu.data(), DIN * 4, hipMemcpyHostToDevice, st));
    GTS_HIP(hipMemcpyAsync(m.sd, sd.data(), DIN * 4, hipMemcpyHostToDevice, st));

    // weights: He-style initialisation, FP32 masters + BF16 copy
    std::mt19937_64 rng(seed);
    std::normal_distribution<float> nd(0.f, 1.f);
    std::vector<float> hw(NW, 0.f);
    for (int i = 0; i < DH * DIN; i++) hw[OW1 + i] = nd(rng) * std::sqrt(2.0f / DIN);
    for (int i = 0; i < DH * DH; i++) hw[OW2 + i] = nd(rng) * std::sqrt(2.0f / DH);
    for (int i = 0; i < DH; i++) hw[OW3 + i] = nd(rng) * std::sqrt(1.0f / DH);
    if (m.w) (void)hipFree(m.w);
    if (m.wb) (void)hipFree(m.wb);
    GTS_HIP(hipMalloc((void **)&m.w, NW * 4 * 4));  // w, g, adam m, adam v
    GTS_HIP(hipMalloc((void **)&m.wb, NW * 2));
    m.nw = NW;
    float *w = m.w, *g = m.w + NW, *am = m.w + 2 * NW, *av = m.w + 3 * NW;
    GTS_HIP(hipMemcpyAsync(w, hw.data(), NW * 4, hipMemcpyHostToDevice, st));
    GTS_HIP(hipMemsetAsync(am, 0, 2 * NW * 4, st));
    k_to_bf16<<<grid_for(NW, 256), 256, 0, st>>>(w, m.wb, NW);

    // training buffers
    const int64_t B = std::min<int64_t>(ntrain, 2048);
    Scratch sx, sbuf;
    float *X = sx.get(ntrain * (DIN + 1));
    float *T = X + ntrain * DIN;
    GTS_HIP(hipMemcpyAsync(X, hf.data(), ntrain * DIN * 4, hipMemcpyHostToDevice, st));
    GTS_HIP(hipMemcpyAsync(T, ty.data(), ntrain * 4, hipMemcpyHostToDevice, st));
    float *buf = sbuf.get((size_t)B * (6 * DH + 2) + 64);
    float *z1 = buf, *a1 = z1 + B * DH, *z2 = a1 + B * DH, *a2 = z2 + B * DH, *d1 = a2 + B * DH,
          *d2 = d1 + B * DH, *y = d2 + B * DH, *e = y + B, *loss = e + B;
    int64_t step = 0;
    float last = 0.f;
    for (int ep = 0; ep < epochs; ep++) {
        GTS_HIP(hipMemsetAsync(loss, 0, 4, st));
        int nbatch = 0;
        for (int64_t s0 = 0; s0 < ntrain; s0 += B) {
            const int64_t nb = std::min<int64_t>(B, ntrain - s0);
            k_mlp_fwd<<<grid_for(nb, 256), 256, 0, st>>>(X + s0 * DIN, nb, m.wb, z1, a1, z2, a2, y);
            k_mlp_bwd<<<grid_for(nb, 256), 256, 0, st>>>(T + s0, nb, m.wb, z1, z2, y, e, d1, d2, loss);
            k_mlp_grad<<<grid_for(NW, 64), 64, 0, st>>>(X + s0 * DIN, nb, a1, a2, e, d1, d2, g);
            step++;
            const float c1 = 1.0f - std::pow(0.9f, (float)step), c2 = 1.0f - std::pow(0.999f, (float)step);
            k_adam<<<grid_for(NW, 256), 256, 0, st>>>(w, g, am, av, m.wb, lr, c1, c2);
            nbatch++;
        }
        GTS_HIP(hipGetLastError());
        if (ep == epochs - 1) {
            GTS_HIP(hipMemcpyAsync(&last, loss, 4, hipMemcpyDeviceToHost, st));
            GTS_HIP(hipStreamSynchronize(st));
            last /= nbatch;
        }
    }

    // safety factor: the 99th percentile of log(true / predicted) on the training set
    float *yp = sbuf.get((size_t)ntrain + 64);
    k_mlp_infer<<<grid_for(ntrain, 4), 256, 0, st>>>(X, ntrain, m.wb, nullptr, 0, 0.f, 1.f, 1.f, yp, nullptr);
    GTS_HIP(hipGetLastError());
    std::vector<float> hy(ntrain);
    GTS_HIP(hipMemcpyAsync(hy.data(), yp, ntrain * 4, hipMemcpyDeviceToHost, st));
    GTS_HIP(hipStreamSynchronize(st));
    std::vector<float> res(ntrain);
    for (int64_t i = 0; i < ntrain; i++) res[i] = (float)((ty[i] - hy[i]) * ys);
    std::sort(res.begin(), res.end());
    m.resid = res;
    learned_apply_quantile(m);
    m.kmax = kmax;
    m.trained = true;
    return last;
}

// safety factor: the chosen quantile of log(true / predicted) on the training
// set (1: the largest), never below 1
void learned_apply_quantile(LearnedModel &m)
{
    const int64_t n = (int64_t)m.resid.size();
    if (n == 0) return;
    const float q = m.resid[std::min<int64_t>(n - 1, (int64_t)((double)m.quantile * n))];
    m.scale = std::exp(std::max(q, 0.0f)) * 1.001f;
}

bool learned_predict(gts_index *I, const int32_t *qpos, const int32_t *kq, int64_t nq, float *bnd)
{
    if (I->radius_src == GTS_RADIUS_PIVOTS) {
        pivot_radius(I, qpos, kq, nq, bnd);
        return false;
    }
    LearnedModel &m = I->model;
    if (!m.trained) fail(GTS_E_ARG, "the radius network is not trained");
    hipStream_t st = I->st;
    constexpr int FR = 16;
    if (m.phi && nq <= FEAT_CH && m.npiv <= 64 * FR) {
        // one table, then features and network fused (the query image staged
        // for the table stays valid for qpos[0..nq))
        float *D = I->ftab.ensure(nq * m.npiv);
        pivot_table_bf16(I, qpos, nq, m.phi, m.pn2, m.npiv, D);
        int ncu = 0;
        GTS_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, I->device));
        const unsigned grid = (unsigned)std::min<int64_t>((nq + 3) / 4, 4 * (int64_t)std::max(ncu, 1));
        k_feat_mlp<FR><<<grid, 256, 0, st>>>(D, nq, m.npiv, m.mu, m.sd, m.wb, kq, m.kmax, m.y_mu, m.y_sd, m.scale,
                                            bnd);
        GTS_HIP(hipGetLastError());
        return true;
    }
    float *feat = I->dfeat.ensure(nq * DIN);
    float *y = I->dpred.ensure(nq);
    features(I, qpos, nq, m.mu, m.sd, feat);
    k_mlp_infer<<<grid_for(nq, 4), 256, 0, st>>>(feat, nq, m.wb, kq, m.kmax, m.y_mu, m.y_sd, m.scale, y, bnd);
    GTS_HIP(hipGetLastError());
    return false;
}

}  // namespace gts
