// batch3d_t3_f16.hip -- the batched three-sweep-shape 3D kernels for FP16 items:
// FP16 -> FP16 (one pass), FP16 -> FP32 and FP32 -> FP16 (the first and last pass
// of an item whose state stays FP32 between passes).
#include "batch3d_kernel.hpp"

namespace ssam {

const void* batch3d_t3_kernel_fn_f16(int storage, int out_storage, int variant) {
    if (storage == SSAM_F16 && out_storage == SSAM_F16) return batch3d_t3_fn<f16_t, f16_t>(variant);
    if (storage == SSAM_F16 && out_storage == SSAM_F32) return batch3d_t3_fn<f16_t, float>(variant);
    if (storage == SSAM_F32 && out_storage == SSAM_F16) return batch3d_t3_fn<float, f16_t>(variant);
    return nullptr;
}

}  // namespace ssam
