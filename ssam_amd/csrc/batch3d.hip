// batch3d.hip -- the 3D kernels of the batched launch: one sweep of radius-1
// stars / face+edge stencils over a table of (item, strip, y-block, z-segment)
// tiles, from a persistent grid draining the XCD queues or an ordinary grid with
// one tile per wave.
#include "batch_common.hpp"

namespace ssam {

// tile.x item, .y strip, .z y-block, .w first plane (low 20 bits) | planes << 20
// GRID: one tile per wave of an ordinary grid over the table (the hardware's
// workgroup dispatch is the queue: no atomics, no drain loop); otherwise a
// persistent grid draining the XCD queues.
template <typename S, int VAR, int SHAPE, bool GRID>
__global__ __launch_bounds__(kWave * kBatchWavesPerBlock) __attribute__((amdgpu_waves_per_eu(Var3DT<S, VAR>::WPE)))
void batch3d_kernel(const BatchItem3D* __restrict__ items, const int4* __restrict__ tiles, QueueSet qs,
                    int* __restrict__ counters, const S* __restrict__ taps, int step) {
    using T = Var3DT<S, VAR>;
    auto body = [&](int t) {
        const int4 e = tiles[t];
        const int it = __builtin_amdgcn_readfirstlane(e.x);
        const int strip = __builtin_amdgcn_readfirstlane(e.y);
        const int yb = __builtin_amdgcn_readfirstlane(e.z);
        const int zw = __builtin_amdgcn_readfirstlane(e.w);
        const int zs = zw & 0xFFFFF, ze = zs + (zw >> 20);
        const BatchItem3D& b = items[it];
        // ping-pong so the last sweep writes dst (same rule as ssam_stencil3d)
        const void* in = step == 0 ? b.src : (((b.steps - step) % 2 == 0) ? b.dst : b.work);
        void* out = ((b.steps - 1 - step) % 2 == 0) ? b.dst : b.work;
        Stencil3DHL<S, SHAPE, T::V, T::NY, T::PF>::run_tile(static_cast<const S*>(in),
                                                             static_cast<S*>(out), b.g, b.vec != 0,
                                                             taps + b.taps_off, strip, yb, zs, ze);
    };
    if constexpr (GRID) {
        (void)counters;
        const int n = qs.beg[kQueues];
        const int t = blockIdx.x * kBatchWavesPerBlock + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
        if (t < n) body(qs.alternate && (step & 1) ? n - 1 - t : t);
    } else {
        drain_queues<false>(qs, counters, qs.alternate && (step & 1), body);
    }
}

template <typename S>
using Batch3DFn = void (*)(const BatchItem3D*, const int4*, QueueSet, int*, const S*, int);

template <typename S, int SHAPE, bool GRID> Batch3DFn<S> batch3d_fn_shape(int variant) {
    switch (variant) {
        case 0: return batch3d_kernel<S, 0, SHAPE, GRID>;
        case 1: return batch3d_kernel<S, 1, SHAPE, GRID>;
        case 2: return batch3d_kernel<S, 2, SHAPE, GRID>;
        case 3: return batch3d_kernel<S, 3, SHAPE, GRID>;
        case 4: return batch3d_kernel<S, 4, SHAPE, GRID>;
        case 5: return batch3d_kernel<S, 5, SHAPE, GRID>;
        case 6: return batch3d_kernel<S, 6, SHAPE, GRID>;
        default: return batch3d_kernel<S, 7, SHAPE, GRID>;
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

const void* batch3d_kernel_fn(int storage, int variant, int shape3d, bool grid) {
    return storage == SSAM_F64 ? batch3d_fn_t<double>(variant, shape3d, grid)
                               : batch3d_fn_t<float>(variant, shape3d, grid);
}

}  // namespace ssam
