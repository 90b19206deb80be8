// batch3d_t2.hip -- the batched 3D kernel with two sweeps per pass (temporal
// blocking, Stencil3DHL2), always as an ordinary grid with one tile per wave.
#include "batch3d_kernel.hpp"

namespace ssam {

const void* batch3d_t2_kernel_fn(int storage, int shape3d) {
    if (storage == SSAM_F64)
        return shape3d ? reinterpret_cast<const void*>(batch3d_kernel<double, kVarT2, k3FaceEdge, true>)
                       : reinterpret_cast<const void*>(batch3d_kernel<double, kVarT2, k3Star, true>);
    return shape3d ? reinterpret_cast<const void*>(batch3d_kernel<float, kVarT2, k3FaceEdge, true>)
                   : reinterpret_cast<const void*>(batch3d_kernel<float, kVarT2, k3Star, true>);
}

}  // namespace ssam
