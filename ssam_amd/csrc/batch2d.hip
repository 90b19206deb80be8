// batch2d.hip -- the 2D kernels of the batched launch: one persistent grid
// draining the XCD tile queues, or an ordinary grid with one tile per wave,
// per launch kind (or one fused kernel for all kinds).
#include "batch_common.hpp"

namespace ssam {

// One tile of one launch kind.
template <typename S, int KIND>
__device__ __forceinline__ void run_tile2d(const S* src, S* dst, const BatchItem2D& b,
                                           const typename Storage<S>::acc* w, int strip, int seg) {
    if constexpr (KIND == k2Star1) Inst2D<S, 1, kStar>::K::run_tile(src, dst, b.p, w, strip, seg);
    else if constexpr (KIND == k2Box1) Inst2D<S, 1, kBox>::K::run_tile(src, dst, b.p, w, strip, seg);
    else if constexpr (KIND == k2Box2) Inst2D<S, 2, kBox>::K::run_tile(src, dst, b.p, w, strip, seg);
    else if constexpr (KIND == k2Box5) Inst2D<S, 5, kBox>::K::run_tile(src, dst, b.p, w, strip, seg);
    else if constexpr (kHalf2D<S> && KIND == k2Box2M)
        Mfma2D<S, 2, std::is_same<S, f16_t>::value, 2, 1>::run_tile(src, dst, b.pm, w, strip, seg);
    else if constexpr (kHalf2D<S> && KIND == k2Box5M)
        Mfma2D<S, 5, std::is_same<S, f16_t>::value, 2, 1>::run_tile(src, dst, b.pm, w, strip, seg);
}

// Waves per SIMD the register allocation leaves room for: the matrix-core tiles
// as their single-grid launch (3 with one set of taps, 2 with hi + lo taps), the
// rest 2 (the fused form holds the largest tile body of all kinds).
template <typename S, int KIND> struct Waves2D {
    static constexpr int value =
        (KIND == k2Box2M || KIND == k2Box5M) ? (std::is_same<S, f16_t>::value ? 2 : 3) : 2;
};

// KIND >= 0: every tile of the launch has that kind (the split form, one launch
// per kind); KIND < 0: the fused form, one launch for all kinds, dispatching per
// item.
// GRID: one tile per wave of an ordinary grid over the table (the hardware's
// workgroup dispatch is the queue); otherwise a persistent grid draining the XCD
// queues.
template <typename S, int KIND, bool GRID>
__global__ __launch_bounds__(kWave * kBatchWavesPerBlock)
__attribute__((amdgpu_waves_per_eu(Waves2D<S, KIND>::value)))
void batch2d_kernel(const BatchItem2D* __restrict__ items, const int4* __restrict__ tiles, QueueSet qs,
                    int* __restrict__ counters, const typename Storage<S>::acc* __restrict__ taps) {
    auto body = [&](int t) {
        const int4 e = tiles[t];
        const int it = __builtin_amdgcn_readfirstlane(e.x);
        const int strip = __builtin_amdgcn_readfirstlane(e.y);
        const int seg = __builtin_amdgcn_readfirstlane(e.z);
        const BatchItem2D& b = items[it];
        const S* src = static_cast<const S*>(b.src);
        S* dst = static_cast<S*>(b.dst);
        const typename Storage<S>::acc* w = taps + b.taps_off;
        if constexpr (KIND >= 0) {
            run_tile2d<S, KIND>(src, dst, b, w, strip, seg);
        } else {
            switch (b.kind) {
                case k2Star1: run_tile2d<S, k2Star1>(src, dst, b, w, strip, seg); break;
                case k2Box1: run_tile2d<S, k2Box1>(src, dst, b, w, strip, seg); break;
                case k2Box2: run_tile2d<S, k2Box2>(src, dst, b, w, strip, seg); break;
                case k2Box2M: run_tile2d<S, k2Box2M>(src, dst, b, w, strip, seg); break;
                case k2Box5M: run_tile2d<S, k2Box5M>(src, dst, b, w, strip, seg); break;
                default: run_tile2d<S, k2Box5>(src, dst, b, w, strip, seg); break;
            }
        }
    };
    if constexpr (GRID) {
        (void)counters;
        const int t = blockIdx.x * kBatchWavesPerBlock + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
        if (t < qs.beg[kQueues]) body(t);
    } else {
        drain_queues<true>(qs, counters, false, body);
    }
}


template <typename S, bool GRID> const void* batch2d_fn_s(int kind) {
    switch (kind) {
        case k2Star1: return reinterpret_cast<const void*>(batch2d_kernel<S, k2Star1, GRID>);
        case k2Box1: return reinterpret_cast<const void*>(batch2d_kernel<S, k2Box1, GRID>);
        case k2Box2: return reinterpret_cast<const void*>(batch2d_kernel<S, k2Box2, GRID>);
        case k2Box5: return reinterpret_cast<const void*>(batch2d_kernel<S, k2Box5, GRID>);
        case k2Box2M: return reinterpret_cast<const void*>(batch2d_kernel<S, k2Box2M, GRID>);
        case k2Box5M: return reinterpret_cast<const void*>(batch2d_kernel<S, k2Box5M, GRID>);
        default: return reinterpret_cast<const void*>(batch2d_kernel<S, -1, GRID>);
    }
}

template <typename S> const void* batch2d_fn_g(int kind, bool grid) {
    return grid ? batch2d_fn_s<S, true>(kind) : batch2d_fn_s<S, false>(kind);
}

const void* batch2d_kernel_fn(int storage, int kind, bool grid) {
    switch (storage) {
        case SSAM_F32: return batch2d_fn_g<float>(kind, grid);
        case SSAM_F64: return batch2d_fn_g<double>(kind, grid);
        case SSAM_BF16: return batch2d_fn_g<bf16_t>(kind, grid);
        default: return batch2d_fn_g<f16_t>(kind, grid);
    }
}

}  // namespace ssam
