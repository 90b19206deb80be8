// batch3d_t2.hip -- the batched 3D kernel with two sweeps per pass (temporal
// blocking, Stencil3DHL2), always as an ordinary grid with one tile per wave.
#include "batch3d_kernel.hpp"

namespace ssam {

const void* batch3d_t2_kernel_fn(int storage, int out_storage, int shape3d) {
    if (storage == SSAM_BF16 || out_storage == SSAM_BF16) return batch3d_t2_kernel_fn_bf16(storage, out_storage, shape3d);
    if (storage == SSAM_F16 || out_storage == SSAM_F16) return batch3d_t2_kernel_fn_f16(storage, out_storage, shape3d);
    if (storage != out_storage) return nullptr;
    if (storage == SSAM_F64) return batch3d_t2_fn<double, double>(shape3d);
    return batch3d_t2_fn<float, float>(shape3d);
}

}  // namespace ssam
