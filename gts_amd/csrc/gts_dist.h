// Device distance helpers for gfx950 (64-lane wavefronts).
//
// A "lane group" of LPO lanes computes one distance: each lane walks the
// dimensions with stride LPO (coalesced across the group) and the partial
// results are combined with xor-shuffles inside the group.  Metrics follow the
// reference's definitions (tree.cuh:81-127): L2 = sqrt(sum (a-b)^2),
// L1 = sum |a-b|, Linf = max |a-b|, angular = |acos(a.b / (|a| |b|))| * 180 /
// 3.1415926 degrees with the cosine clamped to [-1, 1].
#pragma once
#include <hip/hip_runtime.h>

#include "gts_edit.h"
#include "gts_internal.h"

namespace gts {

template <int M>
struct Acc;

template <>
struct Acc<GTS_METRIC_L2> {
    static __device__ __forceinline__ float step(float acc, float a, float b)
    {
        float d = a - b;
        return fmaf(d, d, acc);
    }
    static __device__ __forceinline__ float comb(float x, float y) { return x + y; }
    static __device__ __forceinline__ float fin(float x) { return sqrtf(x); }
    static __device__ __forceinline__ double step64(double acc, float a, float b)
    {
        double d = (double)a - (double)b;
        return fma(d, d, acc);
    }
    static __device__ __forceinline__ double comb64(double x, double y) { return x + y; }
    static __device__ __forceinline__ double fin64(double x) { return sqrt(x); }
};

template <>
struct Acc<GTS_METRIC_L1> {
    static __device__ __forceinline__ float step(float acc, float a, float b)
    {
        return acc + fabsf(a - b);
    }
    static __device__ __forceinline__ float comb(float x, float y) { return x + y; }
    static __device__ __forceinline__ float fin(float x) { return x; }
    static __device__ __forceinline__ double step64(double acc, float a, float b)
    {
        return acc + fabs((double)a - (double)b);
    }
    static __device__ __forceinline__ double comb64(double x, double y) { return x + y; }
    static __device__ __forceinline__ double fin64(double x) { return x; }
};

template <>
struct Acc<GTS_METRIC_LINF> {
    static __device__ __forceinline__ float step(float acc, float a, float b)
    {
        return fmaxf(acc, fabsf(a - b));
    }
    static __device__ __forceinline__ float comb(float x, float y) { return fmaxf(x, y); }
    static __device__ __forceinline__ float fin(float x) { return x; }
    static __device__ __forceinline__ double step64(double acc, float a, float b)
    {
        return fmax(acc, fabs((double)a - (double)b));
    }
    static __device__ __forceinline__ double comb64(double x, double y) { return fmax(x, y); }
    static __device__ __forceinline__ double fin64(double x) { return x; }
};

// Angular distance in degrees from a dot product and the two squared norms.
__device__ __forceinline__ float angle_deg(float dot, float na2, float nb2)
{
    float c = dot / (sqrtf(na2) * sqrtf(nb2));
    c = fminf(fmaxf(c, -1.0f), 1.0f);
    return fabsf(acosf(c) * (float)(180.0 / 3.1415926));
}
__device__ __forceinline__ double angle_deg64(double dot, double na2, double nb2)
{
    double c = dot / (sqrt(na2) * sqrt(nb2));
    c = fmin(fmax(c, -1.0), 1.0);
    return fabs(acos(c) * 180.0 / 3.1415926);
}

// Angular: tiles accumulate the dot product; the angle needs both norms.
template <>
struct Acc<GTS_METRIC_ANGULAR> {
    static __device__ __forceinline__ float step(float acc, float a, float b) { return fmaf(a, b, acc); }
    static __device__ __forceinline__ float comb(float x, float y) { return x + y; }
    static __device__ __forceinline__ double step64(double acc, float a, float b)
    {
        return fma((double)a, (double)b, acc);
    }
    static __device__ __forceinline__ double comb64(double x, double y) { return x + y; }
};

// Final distance of an accumulated tile value for the rows at positions qp
// and xp of the tree (the angular metric reads their squared norms; the
// identical object is at distance 0, search_v2.cuh:338-340).
template <int M>
__device__ __forceinline__ float tile_fin(const DevTree &t, float acc, int64_t qp, int64_t xp)
{
    if constexpr (M == GTS_METRIC_ANGULAR) return qp == xp ? 0.0f : angle_deg(acc, t.xn2[qp], t.xn2[xp]);
    else return Acc<M>::fin(acc);
}

// FP32 distance of rows a, b (dim floats) by a group of LPO lanes; all lanes
// of the group receive the result.
template <int M, int LPO>
__device__ __forceinline__ float group_dist(const float *__restrict__ a,
                                            const float *__restrict__ b, int dim, int gl)
{
    if constexpr (M == GTS_METRIC_ANGULAR) {
        float dot = 0.0f, na = 0.0f, nb = 0.0f;
        for (int j = gl; j < dim; j += LPO) {
            const float x = a[j], y = b[j];
            dot = fmaf(x, y, dot);
            na = fmaf(x, x, na);
            nb = fmaf(y, y, nb);
        }
#pragma unroll
        for (int o = LPO / 2; o > 0; o >>= 1) {
            dot += __shfl_xor(dot, o, LPO);
            na += __shfl_xor(na, o, LPO);
            nb += __shfl_xor(nb, o, LPO);
        }
        return angle_deg(dot, na, nb);
    } else {
        float acc = 0.0f;
        for (int j = gl; j < dim; j += LPO) acc = Acc<M>::step(acc, a[j], b[j]);
#pragma unroll
        for (int o = LPO / 2; o > 0; o >>= 1) acc = Acc<M>::comb(acc, __shfl_xor(acc, o, LPO));
        return Acc<M>::fin(acc);
    }
}

// FP64 distance (the re-decision of pairs inside the FP32 error band).
template <int M, int LPO>
__device__ __forceinline__ double group_dist64(const float *__restrict__ a,
                                               const float *__restrict__ b, int dim, int gl)
{
    if constexpr (M == GTS_METRIC_ANGULAR) {
        double dot = 0.0, na = 0.0, nb = 0.0;
        for (int j = gl; j < dim; j += LPO) {
            const double x = a[j], y = b[j];
            dot = fma(x, y, dot);
            na = fma(x, x, na);
            nb = fma(y, y, nb);
        }
#pragma unroll
        for (int o = LPO / 2; o > 0; o >>= 1) {
            dot += __shfl_xor(dot, o, LPO);
            na += __shfl_xor(na, o, LPO);
            nb += __shfl_xor(nb, o, LPO);
        }
        return angle_deg64(dot, na, nb);
    } else {
        double acc = 0.0;
        for (int j = gl; j < dim; j += LPO) acc = Acc<M>::step64(acc, a[j], b[j]);
#pragma unroll
        for (int o = LPO / 2; o > 0; o >>= 1) acc = Acc<M>::comb64(acc, __shfl_xor(acc, o, LPO));
        return Acc<M>::fin64(acc);
    }
}

// The same with 16-byte loads: lane gl of a group of LPO lanes takes the
// float4 groups gl, gl + LPO, ... of the dp-wide rows (dp a multiple of 4,
// 16-byte aligned rows, padding zero in both, dp >= 4), five of them per row
// in flight at a time (a row of dp <= 20 LPO floats in one round of loads).
// Groups past the row end count as zeros, which leave every metric's sum
// unchanged, so the per-lane order of the terms, and the result, match a
// plain loop.
template <int M, int LPO>
__device__ __forceinline__ double group_dist64_v4(const float *__restrict__ a, const float *__restrict__ b, int dp,
                                                  int gl)
{
    constexpr int B = 5;
    const float4 *a4 = (const float4 *)a, *b4 = (const float4 *)b;
    const int n4 = dp >> 2;
    const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    if constexpr (M == GTS_METRIC_ANGULAR) {
        double dot = 0.0, na = 0.0, nb = 0.0;
        for (int j0 = gl; j0 < n4; j0 += B * LPO) {
            float4 xs[B], ys[B];
#pragma unroll
            for (int u = 0; u < B; u++) {
                // unconditional loads (clamped index), zeros selected after
                const int j = j0 + u * LPO, jc = min(j, n4 - 1);
                const float4 xa = a4[jc], yb = b4[jc];
                xs[u] = j < n4 ? xa : z;
                ys[u] = j < n4 ? yb : z;
            }
#pragma unroll
            for (int u = 0; u < B; u++) {
                const float xa[4] = {xs[u].x, xs[u].y, xs[u].z, xs[u].w};
                const float ya[4] = {ys[u].x, ys[u].y, ys[u].z, ys[u].w};
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    dot = fma((double)xa[e], (double)ya[e], dot);
                    na = fma((double)xa[e], (double)xa[e], na);
                    nb = fma((double)ya[e], (double)ya[e], nb);
                }
            }
        }
#pragma unroll
        for (int o = LPO / 2; o > 0; o >>= 1) {
            dot += __shfl_xor(dot, o, LPO);
            na += __shfl_xor(na, o, LPO);
            nb += __shfl_xor(nb, o, LPO);
        }
        return angle_deg64(dot, na, nb);
    } else {
        double acc = 0.0;
        for (int j0 = gl; j0 < n4; j0 += B * LPO) {
            float4 xs[B], ys[B];
#pragma unroll
            for (int u = 0; u < B; u++) {
                // unconditional loads (clamped index), zeros selected after
                const int j = j0 + u * LPO, jc = min(j, n4 - 1);
                const float4 xa = a4[jc], yb = b4[jc];
                xs[u] = j < n4 ? xa : z;
                ys[u] = j < n4 ? yb : z;
            }
#pragma unroll
            for (int u = 0; u < B; u++) {
                acc = Acc<M>::step64(acc, xs[u].x, ys[u].x);
                acc = Acc<M>::step64(acc, xs[u].y, ys[u].y);
                acc = Acc<M>::step64(acc, xs[u].z, ys[u].z);
                acc = Acc<M>::step64(acc, xs[u].w, ys[u].w);
            }
        }
#pragma unroll
        for (int o = LPO / 2; o > 0; o >>= 1) acc = Acc<M>::comb64(acc, __shfl_xor(acc, o, LPO));
        return Acc<M>::fin64(acc);
    }
}

// Dispatch helper: call f.template operator()<M>() for the runtime metric.
#define GTS_METRIC_SWITCH(metric, ...)                                                    \
    switch (metric) {                                                                     \
    case GTS_METRIC_L2: { constexpr int M_ = GTS_METRIC_L2; __VA_ARGS__; } break;         \
    case GTS_METRIC_L1: { constexpr int M_ = GTS_METRIC_L1; __VA_ARGS__; } break;         \
    case GTS_METRIC_LINF: { constexpr int M_ = GTS_METRIC_LINF; __VA_ARGS__; } break;     \
    case GTS_METRIC_ANGULAR: { constexpr int M_ = GTS_METRIC_ANGULAR; __VA_ARGS__; } break; \
    default: ::gts::fail(GTS_E_ARG, "metric not supported by this kernel");               \
    }

}  // namespace gts
