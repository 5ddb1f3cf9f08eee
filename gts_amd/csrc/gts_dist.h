// Device distance helpers for gfx950 (64-lane wavefronts).
//
// A "lane group" of LPO lanes computes one distance: each lane walks the
// dimensions with stride LPO (coalesced across the group) and the partial
// results are combined with xor-shuffles inside the group.  Metrics follow the
// reference's definitions (tree.cuh:81-105): L2 = sqrt(sum (a-b)^2),
// L1 = sum |a-b|, Linf = max |a-b|.
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

// FP32 distance of rows a, b (dim floats) by a group of LPO lanes; all lanes
// of the group receive the result.
template <int M, int LPO>
__device__ __forceinline__ float group_dist(const float *__restrict__ a,
                                            const float *__restrict__ b, int dim, int gl)
{
    float acc = 0.0f;
    for (int j = gl; j < dim; j += LPO) acc = Acc<M>::step(acc, a[j], b[j]);
#pragma unroll
    for (int o = LPO / 2; o > 0; o >>= 1) acc = Acc<M>::comb(acc, __shfl_xor(acc, o, LPO));
    return Acc<M>::fin(acc);
}

// FP64 distance (the re-decision of pairs inside the FP32 error band).
template <int M, int LPO>
__device__ __forceinline__ double group_dist64(const float *__restrict__ a,
                                               const float *__restrict__ b, int dim, int gl)
{
    double acc = 0.0;
    for (int j = gl; j < dim; j += LPO) acc = Acc<M>::step64(acc, a[j], b[j]);
#pragma unroll
    for (int o = LPO / 2; o > 0; o >>= 1) acc = Acc<M>::comb64(acc, __shfl_xor(acc, o, LPO));
    return Acc<M>::fin64(acc);
}

// Dispatch helper: call f.template operator()<M>() for the runtime metric.
#define GTS_METRIC_SWITCH(metric, ...)                                                    \
    switch (metric) {                                                                     \
    case GTS_METRIC_L2: { constexpr int M_ = GTS_METRIC_L2; __VA_ARGS__; } break;         \
    case GTS_METRIC_L1: { constexpr int M_ = GTS_METRIC_L1; __VA_ARGS__; } break;         \
    case GTS_METRIC_LINF: { constexpr int M_ = GTS_METRIC_LINF; __VA_ARGS__; } break;     \
    default: ::gts::fail(GTS_E_ARG, "metric not supported by this kernel");               \
    }

}  // namespace gts
