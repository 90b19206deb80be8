// batch3d_t3_bf16.hip -- the batched three-sweep-shape 3D kernels for BF16 items:
// BF16 -> BF16 (one pass), BF16 -> FP32 and FP32 -> BF16 (the first and last pass
// of an item whose state stays FP32 between passes).
#include "batch3d_kernel.hpp"

namespace ssam {

const void* batch3d_t3_kernel_fn_bf16(int storage, int out_storage, int variant) {
    if (storage == SSAM_BF16 && out_storage == SSAM_BF16) return batch3d_t3_fn<bf16_t, bf16_t>(variant);
    if (storage == SSAM_BF16 && out_storage == SSAM_F32) return batch3d_t3_fn<bf16_t, float>(variant);
    if (storage == SSAM_F32 && out_storage == SSAM_BF16) return batch3d_t3_fn<float, bf16_t>(variant);
    return nullptr;
}

}  // namespace ssam
