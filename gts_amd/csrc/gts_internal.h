// Internal declarations shared by the GTS host runtime and its gfx950 kernels.
#pragma once
#include <cmath>

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/gts.h"

namespace gts {

struct Error {
    int code;
    std::string msg;
};

#define GTS_HIP(call)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            throw ::gts::Error{GTS_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)}; \
    } while (0)

[[noreturn]] inline void fail(int code, const std::string &msg) { throw Error{code, msg}; }

// Thrown, before any of its results are written, by a sub-batch whose search
// would exceed the memory budget of the running call; the batch splitter
// (gts_query.hip) then halves the sub-batch.
struct BudgetExceeded {};

// Grow-only device buffer owned by an index (scratch reused across calls).
template <class T>
struct DBuf {
    T *p = nullptr;
    size_t cap = 0;
    T *ensure(size_t n)
    {
        if (n <= cap && p) return p;
        if (p) GTS_HIP(hipFree(p));
        p = nullptr;
        cap = 0;
        size_t want = n < 16 ? 16 : n + n / 4;
        if (hipMalloc((void **)&p, want * sizeof(T)) != hipSuccess) {
            p = nullptr;
            fail(GTS_E_OOM, "device allocation of " + std::to_string(want * sizeof(T)) + " bytes");
        }
        cap = want;
        return p;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Error model of the FP32 distance kernels: |d_computed - d_exact| <= rel*d + abs.
// Bounds for a sequential FP32 sum of n non-negative terms (any summation tree
// is bounded by the same (n-1)u), doubled for margin:
//   L2   (a-b) rounded (u), squared and accumulated by FMA, sqrt rounded (u)
//   L1   (a-b) rounded (u), |.| exact, accumulated
//   Linf one rounding per difference, max exact
//   edit exact integers
//   angular (degrees) the FP32 cosine dot / (|a| |b|) errs by at most
//        e_c = (4 dim + 32) u (dot, both norms, square roots, division;
//        doubled); |acos x - acos y| <= (pi / sqrt 2) sqrt|x - y| on [-1, 1],
//        so the angle errs by at most K (pi / sqrt 2) sqrt(e_c) degrees
//        (K = 180 / 3.1415926) plus the rounding of acosf and the scaling
struct ErrModel {
    float rel;
    float abs;
};

inline ErrModel err_model(int metric, int dim)
{
    const float u = 5.9604645e-08f;  // 2^-24
    switch (metric) {
    case GTS_METRIC_L2: return {(float)(dim + 8) * u, 1e-30f};
    case GTS_METRIC_L1: return {(float)(2 * dim + 8) * u, 1e-30f};
    case GTS_METRIC_LINF: return {4.0f * u, 1e-30f};
    case GTS_METRIC_ANGULAR:
        return {16.0f * u, (float)(57.295780490442965 * 2.221441469079183 * sqrt((4.0 * dim + 32.0) * (double)u)) +
                               1e-4f};
    default: return {0.0f, 0.0f};
    }
}

struct DevTree {
    const float *X;        // n x dp, rows in tree (position) order, zero padded
    const uint8_t *S;      // strings in tree order, stride sstride
    const int32_t *slen;   // string lengths in tree order
    const int32_t *perm;   // position -> object ID
    const int32_t *nlid;   // node range start
    const int32_t *nsize;  // node range size
    const float *nmin;     // min_dis: d(parent pivot, first object of the node)
    const int32_t *npiv;   // pivot position of a split node, -1 otherwise
    const uint8_t *nflag;  // 1 exists, 2 leaf
    const int32_t *death;  // update path: op index that deleted each position
                           // (INT32_MAX while live), or null
    const int32_t *qop;    // update path: op index of each batch query, or null
                           // (then every recorded deletion is visible)
    const float *ring;     // per node: min_dis of its T children, then +inf (T + 1 floats)
    const uint32_t *cne;   // per node: mask of non-empty children
    const uint32_t *cleaf; // per node: mask of leaf children
    const float *xn2;      // |x|^2 per position (FP32), set once the tree is final
    int64_t n;
    int dim, dp, metric, T, sstride;
};

// Update-path visibility: position pos is hidden from batch query q when it
// was deleted by an op before the query's op (snapshot semantics).
__device__ __forceinline__ bool hidden(const DevTree &t, int64_t pos, int32_t q)
{
    if (!t.death) return false;
    const int32_t d = t.death[pos];
    return d != INT32_MAX && (!t.qop || d < t.qop[q]);
}

struct gts_index_impl;

// device-wide helpers (gts_prims.hip)
void scan_excl_i32_to_i64(const int32_t *in, int64_t *out, int64_t n, int64_t *d_total,
                          hipStream_t st);
void scan_excl_i64(const int64_t *in, int64_t *out, int64_t n, int64_t *d_total, hipStream_t st);
// frees the scans' and the radix sort's workspace of a stream (before the
// stream is destroyed)
void prims_release_stream(hipStream_t st);
// Stable LSD radix sort of (key, value) pairs on the low `bits` bits of the keys.
// Buffers k0/v0 hold the input; the sorted result is returned in *kout/*vout,
// which point at either (k0, v0) or (k1, v1).
void radix_sort_u64(uint64_t *k0, int32_t *v0, uint64_t *k1, int32_t *v1, int64_t n, int bits,
                    uint32_t *hist_scratch, size_t hist_cap, hipStream_t st, uint64_t **kout,
                    int32_t **vout);
size_t radix_hist_words(int64_t n);
void fill_iota_i32(int32_t *p, int64_t n, hipStream_t st);
void fill_i32(int32_t *p, int32_t v, int64_t n, hipStream_t st);
void fill_f32(float *p, float v, int64_t n, hipStream_t st);

}  // namespace gts
