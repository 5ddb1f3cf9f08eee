// BF16 hi+lo MFMA verification form (gts_mfma.hip).
#pragma once

#include "gts_index.h"

namespace gts {

struct Search;
struct VerifyArgs;

// error constants of d^2 for the split-BF16 dot product (see gts_mfma.hip)
struct BfErr {
    float c1, c2;
    float ex = 0.0f;   // E2 also holds 2 |q| ex (FP8 residual rows: the largest row error)
};

// nprod 3: bound of the three-product sums; 1: bound of hi.hi alone (the
// dense kernel's tile gate)
BfErr bf_err_model(int dim, int dk, int nprod);
// keys of the FP32 MFMA tile kernel: E2 = c1 (qn + xn) + 2 c2 |q| |x|, c1 = c2
BfErr fp32_mfma_err_model(int dp);

// Fragment-major layout of the BF16 split copies: for every block of 16 rows
// and every K step of 32 dims, the 64 lanes' 16-byte MFMA B fragments (lane l:
// row l % 16, dims 8 (l / 16) .. +8) are contiguous, so one wave load reads 1
// KB of consecutive memory.  Element offset of (row, dims k .. k+8), k % 8 == 0:
__host__ __device__ __forceinline__ int64_t frag_off(int64_t row, int k, int nks)
{
    return (((row >> 4) * nks + (k >> 5)) * 64 + ((k & 31) >> 3) * 16 + (row & 15)) * 8;
}

// Byte layout of the FP8 residual rows of the cell scan: for every block of
// 32 rows and every K step, lane l's 16 bytes are the 8 dims of row l % 16 of
// the first 16 rows, then those of row 16 + l % 16 (dims 8 (l / 16) .. +8).
// Byte offset of (row, dim):
__host__ __device__ __forceinline__ int64_t frag8_off(int64_t row, int k, int nks)
{
    return (((row >> 5) * nks + (k >> 5)) * 64 + ((k & 31) >> 3) * 16 + (row & 15)) * 16 + ((row >> 4) & 1) * 8 +
           (k & 7);
}

// round-to-nearest-even FP32 -> BF16 bits (finite inputs)
__device__ __forceinline__ uint16_t bf16_rn(float v)
{
    uint32_t u = __float_as_uint(v);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

void ensure_bf16_split(gts_index *idx);
void ensure_xn2_max(gts_index *idx);
void launch_verify_bf16(gts_index *idx, const Search &S, const VerifyArgs &va);

}  // namespace gts
