// batch3d.hip -- the 3D kernels of the batched launch with one sweep per pass:
// radius-1 stars / face+edge stencils over a table of (item, strip, y-block,
// z-segment) tiles, from a persistent grid draining the XCD queues or an
// ordinary grid with one tile per wave.  The two-sweep shape is in batch3d_t2.hip.
#include "batch3d_kernel.hpp"

namespace ssam {


template <typename S, int SHAPE, bool GRID> Batch3DFn<S> batch3d_fn_shape(int variant) {
    switch (variant) {
        case 0: return batch3d_kernel<S, S, 0, SHAPE, GRID>;
        case 1: return batch3d_kernel<S, S, 1, SHAPE, GRID>;
        case 2: return batch3d_kernel<S, S, 2, SHAPE, GRID>;
        case 3: return batch3d_kernel<S, S, 3, SHAPE, GRID>;
        default: return batch3d_kernel<S, S, kVarPair3D, SHAPE, GRID>;
    }
}

template <typename S> const void* batch3d_fn_t(int variant, int shape3d, bool grid) {
    Batch3DFn<S> f;
    if (grid)
        f = shape3d ? batch3d_fn_shape<S, k3FaceEdge, true>(variant) : batch3d_fn_shape<S, k3Star, true>(variant);
    else
        f = shape3d ? batch3d_fn_shape<S, k3FaceEdge, false>(variant) : batch3d_fn_shape<S, k3Star, false>(variant);
    return reinterpret_cast<const void*>(f);
}

const void* batch3d_kernel_fn(int storage, int out_storage, int variant, int shape3d, bool grid) {
    if (kVar3D[variant].t == 3) return shape3d ? nullptr : batch3d_t3_kernel_fn(storage, out_storage, variant);
    if (kVar3D[variant].t == 2) return batch3d_t2_kernel_fn(storage, out_storage, shape3d);
    // one-sweep shapes: F32 and F64 (BF16 / F16 items always run the two-sweep shape)
    if (storage != out_storage) return nullptr;
    if (storage == SSAM_F64) return batch3d_fn_t<double>(variant, shape3d, grid);
    if (storage == SSAM_F32) return batch3d_fn_t<float>(variant, shape3d, grid);
    return nullptr;
}

}  // namespace ssam
