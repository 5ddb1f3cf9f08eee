// BF16 hi+lo MFMA verification form (gts_mfma.hip).
#pragma once

#include "gts_index.h"

namespace gts {

struct Search;
struct VerifyArgs;

// error constants of d^2 for the split-BF16 dot product (see gts_mfma.hip)
struct BfErr {
    float c1, c2;
};

BfErr bf_err_model(int dim, int dk);
void ensure_bf16_split(gts_index *idx);
void launch_verify_bf16(gts_index *idx, const Search &S, const VerifyArgs &va);

}  // namespace gts
