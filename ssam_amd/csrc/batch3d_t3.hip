// batch3d_t3.hip -- the batched 3D kernels of the three-sweep shape for 7-point
// items (Stencil3DHLT; passes of three, two and one sweep), always as an ordinary
// grid with one tile per wave.
#include "batch3d_kernel.hpp"

namespace ssam {

const void* batch3d_t3_kernel_fn(int storage, int out_storage, int variant) {
    if (storage == SSAM_BF16 || out_storage == SSAM_BF16) return batch3d_t3_kernel_fn_bf16(storage, out_storage, variant);
    if (storage == SSAM_F16 || out_storage == SSAM_F16) return batch3d_t3_kernel_fn_f16(storage, out_storage, variant);
    if (storage != out_storage) return nullptr;
    if (storage == SSAM_F64) return batch3d_t3_fn<double, double>(variant);
    return batch3d_t3_fn<float, float>(variant);
}

}  // namespace ssam
