// Device-wide primitives for gfx950: exclusive scans and a stable LSD radix
// sort of 64-bit keys with 32-bit values.  They replace the Thrust calls of
// the reference (reduce / exclusive_scan / sort_by_key, e.g. tree.cuh:372,
// tree.cuh:383, search_v2.cuh:890-891); see the call sites for the semantics.
//
// Radix sort: 8-bit digits; per pass one histogram kernel (LDS atomics per
// 4096-key tile), one exclusive scan of the digit-major histogram, and one
// scatter kernel that ranks keys stably inside the tile with 64-lane ballots
// (peer mask of equal digits per wave) plus per-wave digit counts in LDS.
#include "gts_internal.h"

#include <algorithm>
#include <mutex>
#include <vector>

namespace gts {

namespace {

constexpr int SB = 256;        // scan block
constexpr int SIPT = 8;        // scan items per thread
constexpr int STILE = SB * SIPT;

template <class T>
__device__ inline int64_t block_excl_scan(int64_t v, int64_t *total)
{
    // 256 threads = 4 waves of 64
    __shared__ int64_t wsum[SB / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int64_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        int64_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int64_t pre = 0, tot = 0;
    for (int i = 0; i < SB / 64; i++) {
        if (i < w) pre += wsum[i];
        tot += wsum[i];
    }
    __syncthreads();
    *total = tot;
    return pre + x - v;
}

template <class InT>
__global__ void __launch_bounds__(SB) k_scan_partials(const InT *in, int64_t n, int64_t *part)
{
    int64_t base = (int64_t)blockIdx.x * STILE + (int64_t)threadIdx.x * SIPT;
    int64_t s = 0;
    for (int i = 0; i < SIPT; i++) {
        int64_t idx = base + i;
        if (idx < n) s += (int64_t)in[idx];
    }
    int64_t tot;
    block_excl_scan<int64_t>(s, &tot);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

template <class InT>
__global__ void __launch_bounds__(SB) k_scan_apply(const InT *in, int64_t n, const int64_t *part_ex,
                                                   int64_t *out)
{
    int64_t base = (int64_t)blockIdx.x * STILE + (int64_t)threadIdx.x * SIPT;
    InT v[SIPT];
    int64_t s = 0;
    for (int i = 0; i < SIPT; i++) {
        int64_t idx = base + i;
        v[i] = idx < n ? in[idx] : (InT)0;
        s += (int64_t)v[i];
    }
    int64_t tot;
    int64_t pre = block_excl_scan<int64_t>(s, &tot) + (part_ex ? part_ex[blockIdx.x] : 0);
    for (int i = 0; i < SIPT; i++) {
        int64_t idx = base + i;
        if (idx < n) out[idx] = pre;
        pre += (int64_t)v[i];
    }
}

// single block exclusive scan of up to STILE elements; writes the total
template <class InT>
__global__ void __launch_bounds__(SB) k_scan_small(const InT *in, int64_t n, int64_t *out,
                                                   int64_t *total)
{
    int64_t base = (int64_t)threadIdx.x * SIPT;
    InT v[SIPT];
    int64_t s = 0;
    for (int i = 0; i < SIPT; i++) {
        int64_t idx = base + i;
        v[i] = idx < n ? in[idx] : (InT)0;
        s += (int64_t)v[i];
    }
    int64_t tot;
    int64_t pre = block_excl_scan<int64_t>(s, &tot);
    for (int i = 0; i < SIPT; i++) {
        int64_t idx = base + i;
        if (idx < n) out[idx] = pre;
        pre += (int64_t)v[i];
    }
    if (threadIdx.x == 0 && total) *total = tot;
}

// bytes of partial sums that scan_impl needs below n items
size_t scan_ws_bytes(int64_t n)
{
    size_t b = 0;
    while (n > STILE) {
        const int64_t nb = (n + STILE - 1) / STILE;
        b += sizeof(int64_t) * 2 * (size_t)nb;
        n = nb;
    }
    return b;
}

// ws: scan_ws_bytes(n) bytes of device memory for the partial sums
template <class InT>
void scan_impl(const InT *in, int64_t *out, int64_t n, int64_t *d_total, hipStream_t st, char *ws)
{
    if (n <= 0) {
        if (d_total) GTS_HIP(hipMemsetAsync(d_total, 0, sizeof(int64_t), st));
        return;
    }
    if (n <= STILE) {
        k_scan_small<InT><<<1, SB, 0, st>>>(in, n, out, d_total);
        GTS_HIP(hipGetLastError());
        return;
    }
    int64_t nb = (n + STILE - 1) / STILE;
    int64_t *part = (int64_t *)ws, *part_ex = part + nb;
    k_scan_partials<InT><<<(unsigned)nb, SB, 0, st>>>(in, n, part);
    GTS_HIP(hipGetLastError());
    scan_impl<int64_t>(part, part_ex, nb, d_total, st, ws + sizeof(int64_t) * 2 * nb);
    k_scan_apply<InT><<<(unsigned)nb, SB, 0, st>>>(in, n, part_ex, out);
    GTS_HIP(hipGetLastError());
}

// Workspace of the scans and the radix sort, one grow-only region per stream
// (a stream runs one primitive at a time, and an index owns its stream), so
// that a call allocates nothing once its sizes have been seen.  Growth frees
// the old region in stream order.  Under graph capture the region comes from
// the graph's own stream-ordered allocations instead (a later growth would
// free memory that a captured graph still names).
struct StreamWs {
    hipStream_t st;
    char *p;
    size_t cap;
};
std::mutex g_ws_mu;
std::vector<StreamWs> g_ws;

class Workspace {
  public:
    Workspace(hipStream_t st, size_t bytes) : st_(st)
    {
        if (bytes == 0) return;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        GTS_HIP(hipStreamIsCapturing(st, &cs));
        if (cs != hipStreamCaptureStatusNone) {
            GTS_HIP(hipMallocAsync((void **)&p_, bytes, st));
            owned_ = true;
            return;
        }
        std::lock_guard<std::mutex> lk(g_ws_mu);
        StreamWs *w = nullptr;
        for (StreamWs &x : g_ws)
            if (x.st == st) w = &x;
        if (!w) {
            g_ws.push_back(StreamWs{st, nullptr, 0});
            w = &g_ws.back();
        }
        if (w->cap < bytes) {
            if (w->p) GTS_HIP(hipFreeAsync(w->p, st));
            const size_t cap = std::max(bytes, w->cap * 2);
            w->p = nullptr;
            w->cap = 0;
            GTS_HIP(hipMallocAsync((void **)&w->p, cap, st));
            w->cap = cap;
        }
        p_ = w->p;
    }
    ~Workspace()
    {
        if (owned_) (void)hipFreeAsync(p_, st_);
    }
    char *get() const { return p_; }

  private:
    hipStream_t st_;
    char *p_ = nullptr;
    bool owned_ = false;
};

// ----------------------------------------------------------------- radix sort
constexpr int RB = 256;          // threads
constexpr int RROUNDS = 16;      // rounds of 256 keys per tile
constexpr int RTILE = RB * RROUNDS;

__global__ void __launch_bounds__(RB) k_radix_hist(const uint64_t *keys, int64_t n, int shift,
                                                   int32_t *hist, int nblocks)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    int64_t base = (int64_t)blockIdx.x * RTILE;
    for (int r = 0; r < RROUNDS; r++) {
        int64_t idx = base + r * RB + threadIdx.x;
        if (idx < n) atomicAdd(&h[(unsigned)(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = (int32_t)h[threadIdx.x];
}

__global__ void __launch_bounds__(RB) k_radix_scatter(const uint64_t *kin, const int32_t *vin,
                                                      uint64_t *kout, int32_t *vout, int64_t n,
                                                      int shift, const int64_t *offs, int nblocks)
{
    __shared__ uint32_t running[256];
    __shared__ uint32_t wcnt[RB / 64][256];
    __shared__ int64_t gbase[256];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    running[t] = 0;
    gbase[t] = offs[(int64_t)t * nblocks + blockIdx.x];
    for (int i = 0; i < RB / 64; i++) wcnt[i][t] = 0;
    __syncthreads();
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    int64_t base = (int64_t)blockIdx.x * RTILE;
    for (int r = 0; r < RROUNDS; r++) {
        int64_t idx = base + r * RB + t;
        bool valid = idx < n;
        uint64_t k = valid ? kin[idx] : 0ull;
        int32_t v = valid ? vin[idx] : 0;
        unsigned d = valid ? (unsigned)(k >> shift) & 255u : 0u;
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < 8; b++) {
            bool bit = (d >> b) & 1u;
            uint64_t bal = __ballot(valid && bit);
            peers &= bit ? bal : ~bal;
        }
        unsigned rank = (unsigned)__popcll(peers & lt);
        if (valid && rank == 0) wcnt[w][d] = (unsigned)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pre = running[d];
            for (int i = 0; i < w; i++) pre += wcnt[i][d];
            int64_t pos = gbase[d] + pre + rank;
            kout[pos] = k;
            vout[pos] = v;
        }
        __syncthreads();
        uint32_t add = 0;
        for (int i = 0; i < RB / 64; i++) {
            add += wcnt[i][t];
            wcnt[i][t] = 0;
        }
        running[t] += add;
        __syncthreads();
    }
}

__global__ void k_iota(int32_t *p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = (int32_t)i;
}
__global__ void k_fill_i32(int32_t *p, int32_t v, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_f32(float *p, float v, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

unsigned grid_for(int64_t n, int block)
{
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (unsigned)g;
}

}  // namespace

void scan_excl_i32_to_i64(const int32_t *in, int64_t *out, int64_t n, int64_t *d_total,
                          hipStream_t st)
{
    Workspace ws(st, scan_ws_bytes(n));
    scan_impl<int32_t>(in, out, n, d_total, st, ws.get());
}

void scan_excl_i64(const int64_t *in, int64_t *out, int64_t n, int64_t *d_total, hipStream_t st)
{
    Workspace ws(st, scan_ws_bytes(n));
    scan_impl<int64_t>(in, out, n, d_total, st, ws.get());
}

void prims_release_stream(hipStream_t st)
{
    std::lock_guard<std::mutex> lk(g_ws_mu);
    for (size_t i = 0; i < g_ws.size(); i++)
        if (g_ws[i].st == st) {
            if (g_ws[i].p) (void)hipFree(g_ws[i].p);
            g_ws.erase(g_ws.begin() + i);
            return;
        }
}

size_t radix_hist_words(int64_t n) { return 256 * (size_t)((n + RTILE - 1) / RTILE); }

void radix_sort_u64(uint64_t *k0, int32_t *v0, uint64_t *k1, int32_t *v1, int64_t n, int bits,
                    uint32_t *, size_t, hipStream_t st, uint64_t **kout, int32_t **vout)
{
    uint64_t *ka = k0, *kb = k1;
    int32_t *va = v0, *vb = v1;
    if (n > 1 && bits > 0) {
        int nblocks = (int)((n + RTILE - 1) / RTILE);
        size_t words = 256 * (size_t)nblocks;
        // offsets, histogram and the scan's partial sums in one workspace
        Workspace ws(st, words * (sizeof(int64_t) + sizeof(int32_t)) + scan_ws_bytes((int64_t)words));
        int64_t *offs = (int64_t *)ws.get();
        int32_t *hist = (int32_t *)(offs + words);
        char *sws = (char *)(hist + words);   // words is a multiple of 256: 8-byte aligned
        for (int shift = 0; shift < bits; shift += 8) {
            k_radix_hist<<<nblocks, RB, 0, st>>>(ka, n, shift, hist, nblocks);
            GTS_HIP(hipGetLastError());
            scan_impl<int32_t>(hist, offs, (int64_t)words, nullptr, st, sws);
            k_radix_scatter<<<nblocks, RB, 0, st>>>(ka, va, kb, vb, n, shift, offs, nblocks);
            GTS_HIP(hipGetLastError());
            uint64_t *tk = ka; ka = kb; kb = tk;
            int32_t *tv = va; va = vb; vb = tv;
        }
    }
    *kout = ka;
    *vout = va;
}

void fill_iota_i32(int32_t *p, int64_t n, hipStream_t st)
{
    k_iota<<<grid_for(n, 256), 256, 0, st>>>(p, n);
    GTS_HIP(hipGetLastError());
}
void fill_i32(int32_t *p, int32_t v, int64_t n, hipStream_t st)
{
    k_fill_i32<<<grid_for(n, 256), 256, 0, st>>>(p, v, n);
    GTS_HIP(hipGetLastError());
}
void fill_f32(float *p, float v, int64_t n, hipStream_t st)
{
    k_fill_f32<<<grid_for(n, 256), 256, 0, st>>>(p, v, n);
    GTS_HIP(hipGetLastError());
}

}  // namespace gts
