// Wave-level (64-lane) sorting helpers on DPP, ds_swizzle and ds_bpermute
// lane exchanges (gfx950): a bitonic sort of one 64-bit (or 32-bit) key per
// lane and the merge step of a running top-64.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace gts {

template <int J>
__device__ __forceinline__ uint32_t lane_xor(uint32_t v)
{
    if constexpr (J == 1) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, true);  // quad_perm [1,0,3,2]
    else if constexpr (J == 2) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, true);  // quad_perm [2,3,0,1]
    else if constexpr (J < 32) return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x1F | (J << 10));
    else {
        const int lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0));
        return (uint32_t)__builtin_amdgcn_ds_bpermute((lane ^ 32) << 2, (int)v);
    }
}

template <int J>
__device__ __forceinline__ uint64_t lane_xor64(uint64_t v)
{
    return ((uint64_t)lane_xor<J>((uint32_t)(v >> 32)) << 32) | lane_xor<J>((uint32_t)v);
}

// compare-exchange with the partner lane ^ J: the lower lane keeps the minimum
// when `up`, the maximum otherwise
template <int J>
__device__ __forceinline__ uint64_t cex(uint64_t v, int lane, bool up)
{
    const uint64_t o = lane_xor64<J>(v);
    const bool lower = (lane & J) == 0;
    return (lower == up) ? (v < o ? v : o) : (v < o ? o : v);
}

__device__ __forceinline__ uint64_t wave_sort64(uint64_t v, int lane)
{
    v = cex<1>(v, lane, (lane & 2) == 0);
    v = cex<2>(v, lane, (lane & 4) == 0);
    v = cex<1>(v, lane, (lane & 4) == 0);
    v = cex<4>(v, lane, (lane & 8) == 0);
    v = cex<2>(v, lane, (lane & 8) == 0);
    v = cex<1>(v, lane, (lane & 8) == 0);
    v = cex<8>(v, lane, (lane & 16) == 0);
    v = cex<4>(v, lane, (lane & 16) == 0);
    v = cex<2>(v, lane, (lane & 16) == 0);
    v = cex<1>(v, lane, (lane & 16) == 0);
    v = cex<16>(v, lane, (lane & 32) == 0);
    v = cex<8>(v, lane, (lane & 32) == 0);
    v = cex<4>(v, lane, (lane & 32) == 0);
    v = cex<2>(v, lane, (lane & 32) == 0);
    v = cex<1>(v, lane, (lane & 32) == 0);
    v = cex<32>(v, lane, true);
    v = cex<16>(v, lane, true);
    v = cex<8>(v, lane, true);
    v = cex<4>(v, lane, true);
    v = cex<2>(v, lane, true);
    return cex<1>(v, lane, true);
}

// ascending order of a bitonic sequence across the 64 lanes
__device__ __forceinline__ uint64_t wave_clean64(uint64_t v, int lane)
{
    v = cex<32>(v, lane, true);
    v = cex<16>(v, lane, true);
    v = cex<8>(v, lane, true);
    v = cex<4>(v, lane, true);
    v = cex<2>(v, lane, true);
    return cex<1>(v, lane, true);
}

__device__ __forceinline__ uint64_t lane_read64(uint64_t v, int src)
{
    return ((uint64_t)(uint32_t)__shfl((int)(v >> 32), src, 64) << 32) | (uint32_t)__shfl((int)(uint32_t)v, src, 64);
}

// Key of lane `src` (wave-uniform src), read with readlane.
__device__ __forceinline__ uint64_t lane_key64(uint64_t v, int src)
{
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(v >> 32), src);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, src);
    return ((uint64_t)hi << 32) | lo;
}

// Running top-64 of a stream of keys: merge one chunk (one key per lane) into
// `best` (ascending across the lanes).
__device__ __forceinline__ uint64_t wave_merge64(uint64_t best, uint64_t chunk, int lane)
{
    chunk = wave_sort64(chunk, lane);
    const uint64_t r = lane_read64(chunk, 63 - lane);
    return wave_clean64(best < r ? best : r, lane);
}

// Adds every lane's v to *dst with one global atomic per block: wave sums,
// then a block sum in LDS.  Every thread of the block calls it (it holds two
// barriers).  Same-address atomics from thousands of waves that finish
// together serialise; a block's one does not.
// 32-bit keys (e.g. the bits of non-negative floats): the same network on
// one register per lane
template <int J>
__device__ __forceinline__ uint32_t cex32(uint32_t v, int lane, bool up)
{
    const uint32_t o = lane_xor<J>(v);
    const bool lower = (lane & J) == 0;
    return (lower == up) ? (v < o ? v : o) : (v < o ? o : v);
}

__device__ __forceinline__ uint32_t wave_sort32(uint32_t v, int lane)
{
    v = cex32<1>(v, lane, (lane & 2) == 0);
    v = cex32<2>(v, lane, (lane & 4) == 0);
    v = cex32<1>(v, lane, (lane & 4) == 0);
    v = cex32<4>(v, lane, (lane & 8) == 0);
    v = cex32<2>(v, lane, (lane & 8) == 0);
    v = cex32<1>(v, lane, (lane & 8) == 0);
    v = cex32<8>(v, lane, (lane & 16) == 0);
    v = cex32<4>(v, lane, (lane & 16) == 0);
    v = cex32<2>(v, lane, (lane & 16) == 0);
    v = cex32<1>(v, lane, (lane & 16) == 0);
    v = cex32<16>(v, lane, (lane & 32) == 0);
    v = cex32<8>(v, lane, (lane & 32) == 0);
    v = cex32<4>(v, lane, (lane & 32) == 0);
    v = cex32<2>(v, lane, (lane & 32) == 0);
    v = cex32<1>(v, lane, (lane & 32) == 0);
    v = cex32<32>(v, lane, true);
    v = cex32<16>(v, lane, true);
    v = cex32<8>(v, lane, true);
    v = cex32<4>(v, lane, true);
    v = cex32<2>(v, lane, true);
    return cex32<1>(v, lane, true);
}

// the running top-64 step on 32-bit keys (see wave_merge64)
__device__ __forceinline__ uint32_t wave_merge32(uint32_t best, uint32_t chunk, int lane)
{
    chunk = wave_sort32(chunk, lane);
    const uint32_t r = (uint32_t)__shfl((int)chunk, 63 - lane, 64);
    uint32_t v = best < r ? best : r;
    v = cex32<32>(v, lane, true);
    v = cex32<16>(v, lane, true);
    v = cex32<8>(v, lane, true);
    v = cex32<4>(v, lane, true);
    v = cex32<2>(v, lane, true);
    return cex32<1>(v, lane, true);
}

__device__ __forceinline__ void block_add_u64(unsigned long long v, unsigned long long *dst)
{
    __shared__ unsigned long long s_sum;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (threadIdx.x == 0) s_sum = 0;
    __syncthreads();
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(&s_sum, v);
    __syncthreads();
    if (threadIdx.x == 0 && s_sum) atomicAdd(dst, s_sum);
}

}  // namespace gts
