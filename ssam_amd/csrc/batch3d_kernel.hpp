// batch3d_kernel.hpp -- the batched 3D kernel template (instantiated by
// batch3d.hip for the one-sweep tile shapes, by batch3d_t2*.hip for the two-sweep
// shape and by batch3d_t3*.hip for the three-sweep one, so they compile in
// parallel).
#pragma once
#include "batch_common.hpp"
#include "stencil3d_hlt.hpp"

namespace ssam {

// tile.x item, .y strip, .z y-block, .w first plane (low 20 bits) | planes << 20
// GRID: one tile per wave of an ordinary grid over the table (the hardware's
// workgroup dispatch is the queue: no atomics, no drain loop); otherwise a
// persistent grid draining the XCD queues.
template <typename S, int VAR> struct Waves3D { static constexpr int value = Var3DT<S, VAR>::WPE; };
template <typename S> struct Waves3D<S, kVarPair3D> {
    static constexpr int a = Var3DT<S, kPairA>::WPE, b = Var3DT<S, kPairB>::WPE;
    static constexpr int value = a < b ? a : b;
};

// single: the pass runs one sweep (two-sweep shape, the last pass of an odd
// count).  A three-sweep shape's kernel runs T::NSW sweeps on the tiles of three
// (the same halo lanes).
template <typename SI, typename SO, int SHAPE, int VAR>
__device__ __forceinline__ void run_tile3d(const SI* in, SO* out, const BatchItem3D& b,
                                           const typename Storage<SI>::acc* w, int strip, int yb, int zs, int ze,
                                           int single) {
    using T = Var3DT<SI, VAR>;
    if constexpr (T::T == 3) {
        (void)single;
        constexpr int HL = (3 + T::V - 1) / T::V;
        Stencil3DHLT<SI, SHAPE, T::V, T::NY, T::PF, SO, T::NSW, HL>::run_tile(in, out, b.g, b.vec != 0, w, strip,
                                                                              yb, zs, ze);
    } else if constexpr (T::T == 2) {
        Stencil3DHL2<SI, SHAPE, T::V, T::NY, T::PF, SO>::run_tile(in, out, b.g, b.vec != 0, w, strip, yb, zs, ze,
                                                                   single != 0);
    } else {
        (void)single;
        Stencil3DHL<SI, SHAPE, T::V, T::NY, T::PF, SO>::run_tile(in, out, b.g, b.vec != 0, w, strip, yb, zs, ze);
    }
}

// VAR == kVarPair3D: each item runs the tile shape the plan chose for it.
// SI / SO: storage read and written by this pass (a BF16 / F16 item of several
// passes reads its storage type first, writes it last, and keeps FP32 between).
template <typename SI, typename SO, int VAR, int SHAPE, bool GRID>
__global__ __launch_bounds__(kWave * kBatchWavesPerBlock) __attribute__((amdgpu_waves_per_eu(Waves3D<SI, VAR>::value)))
void batch3d_kernel(const BatchItem3D* __restrict__ items, const int4* __restrict__ tiles, QueueSet qs,
                    int* __restrict__ counters, const typename Storage<SI>::acc* __restrict__ taps, int step,
                    int single) {
    auto body = [&](int t) {
        const int4 e = tiles[t];
        const int it = __builtin_amdgcn_readfirstlane(e.x);
        const int strip = __builtin_amdgcn_readfirstlane(e.y);
        const int yb = __builtin_amdgcn_readfirstlane(e.z);
        const int zw = __builtin_amdgcn_readfirstlane(e.w);
        const int zs = zw & 0xFFFFF, ze = zs + (zw >> 20);
        const BatchItem3D& b = items[it];
        // passes ping-pong so the last pass writes dst (same rule as ssam_stencil3d
        // with passes for sweeps); b.steps holds the item's pass count.  An FP32
        // state alternates between the two FP32 grids of work.
        const void* in;
        void* out;
        if (b.fp32_state) {
            const size_t n = static_cast<size_t>(b.g.nx) * b.g.ny * b.g.nz;
            float* st0 = static_cast<float*>(b.work);
            in = step == 0 ? b.src : static_cast<const void*>(st0 + ((step - 1) % 2) * n);
            out = step == b.steps - 1 ? b.dst : static_cast<void*>(st0 + (step % 2) * n);
        } else {
            in = step == 0 ? b.src : (((b.steps - step) % 2 == 0) ? b.dst : b.work);
            out = ((b.steps - 1 - step) % 2 == 0) ? b.dst : b.work;
        }
        const SI* src = static_cast<const SI*>(in);
        SO* dst = static_cast<SO*>(out);
        const typename Storage<SI>::acc* w = taps + b.taps_off;
        if constexpr (VAR == kVarPair3D) {
            if (b.var == kPairB) run_tile3d<SI, SO, SHAPE, kPairB>(src, dst, b, w, strip, yb, zs, ze, 0);
            else run_tile3d<SI, SO, SHAPE, kPairA>(src, dst, b, w, strip, yb, zs, ze, 0);
        } else {
            run_tile3d<SI, SO, SHAPE, VAR>(src, dst, b, w, strip, yb, zs, ze, single);
        }
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
using Batch3DFn = void (*)(const BatchItem3D*, const int4*, QueueSet, int*, const typename Storage<S>::acc*, int,
                           int);

// The two-sweep kernel of one storage pair and shape (batch3d_t2*.hip).
template <typename SI, typename SO>
const void* batch3d_t3_fn(int variant) {
    if (variant == kVarT3S2) return reinterpret_cast<const void*>(batch3d_kernel<SI, SO, kVarT3S2, k3Star, true>);
    if (variant == kVarT3S1) return reinterpret_cast<const void*>(batch3d_kernel<SI, SO, kVarT3S1, k3Star, true>);
    return reinterpret_cast<const void*>(batch3d_kernel<SI, SO, kVarT3, k3Star, true>);
}

template <typename SI, typename SO>
const void* batch3d_t2_fn(int shape3d) {
    return shape3d ? reinterpret_cast<const void*>(batch3d_kernel<SI, SO, kVarT2, k3FaceEdge, true>)
                   : reinterpret_cast<const void*>(batch3d_kernel<SI, SO, kVarT2, k3Star, true>);
}

}  // namespace ssam
