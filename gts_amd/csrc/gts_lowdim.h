// Persistent kNN launch for low-dimensional vector rows (gts_lowdim.hip).
#pragma once

#include "gts_index.h"

namespace gts {

constexpr int LOWDIM_MAX_DIM = 8;   // padded row width the kernel holds in registers
constexpr int LOWDIM_MAX_K = 64;    // one list entry per lane

// True when a learned kNN batch of this index with rows of stride kstride can
// run as the persistent launch: L2, L1 or Linf rows of padded width <= 8,
// kstride <= 64, and a tree whose depth-first stack fits a wave's LDS share.
bool lowdim_supported(const gts_index *idx, int kstride);

// The launch: queries qpos[0..nq) with k kq[i] and radius bnd[i] (null: none)
// into rows orig[i] of ids/dists/kth; *work must be 0 (a fill kernel clears
// it); verified pairs added to *verified; bpc blocks of 4 waves per CU.
void lowdim_launch(gts_index *idx, const int32_t *qpos, const int32_t *kq, const float *bnd, const int32_t *orig,
                   int64_t nq, int kstride, int32_t *ids, float *dists, float *kth, uint32_t *work,
                   unsigned long long *verified, int bpc);

}  // namespace gts
