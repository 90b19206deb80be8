// batch3d_t3.hip -- the batched 3D kernel with three sweeps per pass for 7-point
// FP32 / FP64 items (Stencil3DHLT), always as an ordinary grid with one tile per
// wave.
#include "batch3d_kernel.hpp"

namespace ssam {

// F32 and F64 groups (BF16 / F16 groups run the two-sweep shape)
const void* batch3d_t3_kernel_fn(int storage, int out_storage) {
    if (storage != out_storage) return nullptr;
    if (storage == SSAM_F64) return batch3d_t3_fn<double, double>();
    return batch3d_t3_fn<float, float>();
}

}  // namespace ssam
