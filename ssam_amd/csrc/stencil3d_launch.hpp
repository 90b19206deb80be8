// stencil3d_launch.hpp -- host-side launch of the 3D kernels for one pair of
// storage types (src, dst): one sweep per launch (ssam3d_kernel), two
// (ssam3d_t2_kernel) or three (ssam3d_t3_kernel).  Each pair's instances live
// in their own compile unit: stencil3d.hip (F32, F64), stencil3d_bf16.hip,
// stencil3d_f16.hip (2-byte storage and its FP32 running state).
#pragma once
#include <hip/hip_runtime.h>

#include "ssam_internal.hpp"
#include "stencil3d.hpp"
#include "stencil3d_hl.hpp"
#include "stencil3d_hlt.hpp"
#include "stencil3d_zp.hpp"

namespace ssam {

// the fixed-order sum of a launch's residual partials (stencil3d.hip)
int launch_sum_partials(const double* part, int n, double* out, int take_sqrt, hipStream_t st);

// Columns per lane of the one-sweep kernel by src storage: one 16-byte row access
// for F32 / F64, 8 bytes for BF16 / F16 (four values, the FP32 ring of a
// 16-byte F32 strip).
template <typename S> struct Lanes3D { static constexpr int V = 4; };
template <> struct Lanes3D<double> { static constexpr int V = 2; };

template <typename S, int SHAPE, int R, int V, typename SO = S>
int launch3d(const Req3D& rq, hipStream_t st) {
    using A = typename Storage<S>::acc;
    constexpr int NY = R == 1 ? 4 : 2;
    constexpr int PF = 1;
    constexpr int NT = ntaps3(SHAPE, R);
    Taps3<A, NT> tp;
    for (int i = 0; i < NT; ++i) tp.w[i] = static_cast<A>(rq.taps[i]);
    Params3D p;
    p.nx = static_cast<int>(rq.nx);
    p.ny = static_cast<int>(rq.ny);
    p.nz = static_cast<int>(rq.nz);
    p.sy = rq.nx;
    p.sz = rq.nx * rq.ny;
    p.nstrips = static_cast<int>((rq.nx + kWave * V - 1) / (kWave * V));
    p.nyb = static_cast<int>((rq.ny + NY - 1) / NY);
    p.vec_ok = (rq.nx % V == 0) &&
               (reinterpret_cast<uintptr_t>(rq.src) % (V * sizeof(S)) == 0) &&
               (reinterpret_cast<uintptr_t>(rq.dst) % (V * sizeof(SO)) == 0);
    int zseg = rq.zseg;
    if (zseg <= 0) {
        const long long target = 4LL * 1024;  // waves
        const long long xy = static_cast<long long>(p.nstrips) * p.nyb;
        long long nzs = (target + xy - 1) / xy;
        if (nzs < 1) nzs = 1;
        zseg = static_cast<int>((rq.nz + nzs - 1) / nzs);
        if (zseg < 4 * R) zseg = 4 * R;
    }
    if (zseg > p.nz) zseg = p.nz;
    p.zseg = zseg;
    const long long nzs = (p.nz + zseg - 1) / zseg;
    const long long ntiles = static_cast<long long>(p.nstrips) * p.nyb * nzs;
    if (ntiles > (1LL << 30)) return SSAM_E_UNSUPPORTED;
    p.ntiles = static_cast<int>(ntiles);
    p.nblocks = (p.ntiles + kWaves3D - 1) / kWaves3D;
    p.partials = rq.partials;
    p.res_scale = rq.res_scale;
    if (rq.partials && rq.partials_cap < p.ntiles) return SSAM_E_INVALID;
    hipLaunchKernelGGL((ssam3d_kernel<S, SHAPE, R, V, NY, PF, SO>), dim3(p.nblocks), dim3(kWave * kWaves3D), 0,
                       st, tp, static_cast<const S*>(rq.src), static_cast<SO*>(rq.dst), p);
    if (hipGetLastError() != hipSuccess) return SSAM_E_HIP;
    if (rq.partials) return launch_sum_partials(rq.partials, p.ntiles, rq.res_out, rq.res_sqrt, st);
    return SSAM_OK;
}

// The z-partial-sum form of a radius-R star (no residual).
template <typename S, int R, int V, int NY, int PF, typename SO = S>
int launch3d_zp(const Req3D& rq, hipStream_t st) {
    using A = typename Storage<S>::acc;
    constexpr int NT = ntaps3(k3Star, R);
    Taps3<A, NT> tp;
    for (int i = 0; i < NT; ++i) tp.w[i] = static_cast<A>(rq.taps[i]);
    Params3D p;
    p.nx = static_cast<int>(rq.nx);
    p.ny = static_cast<int>(rq.ny);
    p.nz = static_cast<int>(rq.nz);
    p.sy = rq.nx;
    p.sz = rq.nx * rq.ny;
    p.nstrips = static_cast<int>((rq.nx + kWave * V - 1) / (kWave * V));
    p.nyb = static_cast<int>((rq.ny + NY - 1) / NY);
    p.vec_ok = (rq.nx % V == 0) &&
               (reinterpret_cast<uintptr_t>(rq.src) % (V * sizeof(S)) == 0) &&
               (reinterpret_cast<uintptr_t>(rq.dst) % (V * sizeof(SO)) == 0);
    int zseg = rq.zseg;
    if (zseg <= 0) {
        // about 4096 waves; each tile re-reads 2R planes, so at least 8R planes
        const long long xy = static_cast<long long>(p.nstrips) * p.nyb;
        long long nzs = (4LL * 1024 + xy - 1) / xy;
        if (nzs < 1) nzs = 1;
        zseg = static_cast<int>((rq.nz + nzs - 1) / nzs);
        if (zseg < 8 * R) zseg = 8 * R;
    }
    if (zseg > p.nz) zseg = p.nz;
    p.zseg = zseg;
    const long long ntiles = static_cast<long long>(p.nstrips) * p.nyb * ((p.nz + zseg - 1) / zseg);
    if (ntiles > (1LL << 30)) return SSAM_E_UNSUPPORTED;
    p.ntiles = static_cast<int>(ntiles);
    p.nblocks = (p.ntiles + kWaves3D - 1) / kWaves3D;
    p.partials = nullptr;
    p.res_scale = 0.0;
    hipLaunchKernelGGL((ssam3d_zp_kernel<S, R, V, NY, PF, SO>), dim3(p.nblocks), dim3(kWave * kWaves3D), 0, st,
                       tp, static_cast<const S*>(rq.src), static_cast<SO*>(rq.dst), p);
    return hipGetLastError() == hipSuccess ? SSAM_OK : SSAM_E_HIP;
}

// Rows per lane and planes in flight of the z-partial-sum form, by src storage
// (512^3 radius-2 star, measured against 2/1 and 4/2 resp. 4/1: FP32 4/1 0.231 ms,
// FP64 4/2 0.435 ms; the plane ring 0.262 / 0.513 ms).
template <typename S> struct ZpShape { static constexpr int NY = 4, PF = 1; };
template <> struct ZpShape<double> { static constexpr int NY = 4, PF = 2; };

// One sweep of the request's shape from S to SO.
template <typename S, typename SO>
int dispatch3d_pair(const Req3D& rq, hipStream_t st);

// ---------------------------------------------------------------- two sweeps per launch
// Temporal blocking for radius-1 stars and the 19-point face+edge stencil on
// grids under 2 GiB: one launch runs two sweeps (Stencil3DHL2) with one HBM read
// and one write of the grid.  Tiles (x-strip, y-block, z-segment) in that order,
// one per wave; blocks remapped so each XCD works on a contiguous range of
// tiles (neighbouring y-blocks share their halo rows in that XCD's L2).
template <typename S, int SHAPE, int V, int NY, int PF, typename SO>
__global__ __launch_bounds__(kWave * kWaves3D) __attribute__((amdgpu_waves_per_eu(2)))
void ssam3d_t2_kernel(Taps3<typename Storage<S>::acc, ntaps3(SHAPE, 1)> tp, const S* __restrict__ src,
                      SO* __restrict__ dst, Grid3D g, int nstrips, int nyb, int zseg, int ntiles, int nblocks,
                      int vec) {
    (void)tp;
    using A = typename Storage<S>::acc;
    const A* __restrict__ w = (const A*)(__builtin_amdgcn_kernarg_segment_ptr());
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    int blk = blockIdx.x;
    const int n8 = (nblocks / 8) * 8;
    if (blk < n8) blk = (blk % 8) * (n8 / 8) + blk / 8;
    const int tile = blk * kWaves3D + wv;
    if (tile >= ntiles) return;
    const int strip = tile % nstrips;
    const int rest = tile / nstrips;
    const int yb = rest % nyb;
    const int zs = (rest / nyb) * zseg;
    const int ze = min(g.nz, zs + zseg);
    Stencil3DHL2<S, SHAPE, V, NY, PF, SO>::run_tile(src, dst, g, vec != 0, w, strip, yb, zs, ze, false);
}

template <typename S, int SHAPE, typename SO = S>
int launch3d_t2(const Req3D& rq, hipStream_t st) {
    using A = typename Storage<S>::acc;
    // FP32-accumulated storage: two columns per lane; FP64: one (two halo lanes per side)
    constexpr int V = sizeof(S) == 8 ? 1 : 2;
    constexpr int NY = 6, PF = 1;
    using K = Stencil3DHL2<S, SHAPE, V, NY, PF, SO>;
    constexpr int NT = ntaps3(SHAPE, 1);
    Taps3<A, NT> tp;
    for (int i = 0; i < NT; ++i) tp.w[i] = static_cast<A>(rq.taps[i]);
    Grid3D g;
    g.nx = static_cast<int>(rq.nx);
    g.ny = static_cast<int>(rq.ny);
    g.nz = static_cast<int>(rq.nz);
    g.sy = rq.nx;
    g.sz = rq.nx * rq.ny;
    const int nstrips = static_cast<int>((rq.nx + K::OW - 1) / K::OW);
    const int nyb = static_cast<int>((rq.ny + NY - 1) / NY);
    int zseg = rq.zseg > 0 ? rq.zseg : 16;
    if (zseg > g.nz) zseg = g.nz;
    const long long ntiles = static_cast<long long>(nstrips) * nyb * ((g.nz + zseg - 1) / zseg);
    if (ntiles > (1LL << 30)) return SSAM_E_UNSUPPORTED;
    const int nblocks = static_cast<int>((ntiles + kWaves3D - 1) / kWaves3D);
    const int vec = (rq.nx % V == 0) && (reinterpret_cast<uintptr_t>(rq.src) % (V * sizeof(S)) == 0) &&
                    (reinterpret_cast<uintptr_t>(rq.dst) % (V * sizeof(SO)) == 0);
    hipLaunchKernelGGL((ssam3d_t2_kernel<S, SHAPE, V, NY, PF, SO>), dim3(nblocks), dim3(kWave * kWaves3D), 0, st,
                       tp, static_cast<const S*>(rq.src), static_cast<SO*>(rq.dst), g, nstrips, nyb, zseg,
                       static_cast<int>(ntiles), nblocks, vec);
    return hipGetLastError() == hipSuccess ? SSAM_OK : SSAM_E_HIP;
}

// ---------------------------------------------------------------- three sweeps per launch
// The 7-point star (radius 1) on grids under 2 GiB: three sweeps per launch
// (Stencil3DHLT<.., 3>), one HBM read and write per three sweeps.  Measured on
// 512^3 FP32, ten sweeps as 3+3+2+2 launches: 1.20 ms against 1.36 ms as five
// two-sweep launches (bit-equal results).  The 19-point shape stays at two per
// launch: its third sweep's extra rows cost more than the bytes saved (2.50 vs
// 2.01 ms).  Shape: two columns per lane (FP64 one), 4 rows per lane, 64-plane
// z-segments (16 planes: 1.27 ms).
template <typename S, int V, int NY, int PF, typename SO>
__global__ __launch_bounds__(kWave * kWaves3D) __attribute__((amdgpu_waves_per_eu(2)))
void ssam3d_t3_kernel(Taps3<typename Storage<S>::acc, ntaps3(k3Star, 1)> tp, const S* __restrict__ src,
                      SO* __restrict__ dst, Grid3D g, int nstrips, int nyb, int zseg, int ntiles, int nblocks,
                      int vec) {
    (void)tp;
    using A = typename Storage<S>::acc;
    const A* __restrict__ w = (const A*)(__builtin_amdgcn_kernarg_segment_ptr());
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    int blk = blockIdx.x;
    const int n8 = (nblocks / 8) * 8;
    if (blk < n8) blk = (blk % 8) * (n8 / 8) + blk / 8;
    const int tile = blk * kWaves3D + wv;
    if (tile >= ntiles) return;
    const int strip = tile % nstrips;
    const int rest = tile / nstrips;
    const int yb = rest % nyb;
    const int zs = (rest / nyb) * zseg;
    const int ze = min(g.nz, zs + zseg);
    Stencil3DHLT<S, k3Star, V, NY, PF, SO, 3>::run_tile(src, dst, g, vec != 0, w, strip, yb, zs, ze);
}

template <typename S, typename SO>
int launch3d_t3(const Req3D& rq, hipStream_t st) {
    using A = typename Storage<S>::acc;
    constexpr int V = sizeof(S) == 8 ? 1 : 2;
    constexpr int NY = 4, PF = 1;
    using K = Stencil3DHLT<S, k3Star, V, NY, PF, SO, 3>;
    constexpr int NT = ntaps3(k3Star, 1);
    Taps3<A, NT> tp;
    for (int i = 0; i < NT; ++i) tp.w[i] = static_cast<A>(rq.taps[i]);
    Grid3D g;
    g.nx = static_cast<int>(rq.nx);
    g.ny = static_cast<int>(rq.ny);
    g.nz = static_cast<int>(rq.nz);
    g.sy = rq.nx;
    g.sz = rq.nx * rq.ny;
    const int nstrips = static_cast<int>((rq.nx + K::OW - 1) / K::OW);
    const int nyb = static_cast<int>((rq.ny + NY - 1) / NY);
    int zseg = rq.zseg > 0 ? rq.zseg : 64;
    if (zseg > g.nz) zseg = g.nz;
    const long long ntiles = static_cast<long long>(nstrips) * nyb * ((g.nz + zseg - 1) / zseg);
    if (ntiles > (1LL << 30)) return SSAM_E_UNSUPPORTED;
    const int nblocks = static_cast<int>((ntiles + kWaves3D - 1) / kWaves3D);
    const int vec = (rq.nx % V == 0) && (reinterpret_cast<uintptr_t>(rq.src) % (V * sizeof(S)) == 0) &&
                    (reinterpret_cast<uintptr_t>(rq.dst) % (V * sizeof(SO)) == 0);
    hipLaunchKernelGGL((ssam3d_t3_kernel<S, V, NY, PF, SO>), dim3(nblocks), dim3(kWave * kWaves3D), 0, st, tp,
                       static_cast<const S*>(rq.src), static_cast<SO*>(rq.dst), g, nstrips, nyb, zseg,
                       static_cast<int>(ntiles), nblocks, vec);
    return hipGetLastError() == hipSuccess ? SSAM_OK : SSAM_E_HIP;
}

// Two sweeps of the request's radius-1 shape from S to SO (three: rq.sweeps == 3,
// 7-point star).
template <typename S, typename SO>
int dispatch3d_t2_pair(const Req3D& rq, hipStream_t st);

#ifdef SSAM_3D_PAIR_UNIT
template <typename S, typename SO>
int dispatch3d_pair(const Req3D& rq, hipStream_t st) {
    constexpr int V = Lanes3D<S>::V;
    if (rq.shape == k3FaceEdge) {
        if (rq.radius != 1) return SSAM_E_UNSUPPORTED;
        return launch3d<S, k3FaceEdge, 1, V, SO>(rq, st);
    }
    if (rq.radius == 1) return launch3d<S, k3Star, 1, V, SO>(rq, st);
    if (rq.radius == 2) {
        if (rq.form != 1 && !rq.partials) return launch3d_zp<S, 2, V, ZpShape<S>::NY, ZpShape<S>::PF, SO>(rq, st);
        return launch3d<S, k3Star, 2, V, SO>(rq, st);
    }
    return SSAM_E_UNSUPPORTED;
}

template <typename S, typename SO>
int dispatch3d_t2_pair(const Req3D& rq, hipStream_t st) {
    if (rq.sweeps == 3) return rq.shape == k3Star ? launch3d_t3<S, SO>(rq, st) : SSAM_E_UNSUPPORTED;
    return rq.shape == k3FaceEdge ? launch3d_t2<S, k3FaceEdge, SO>(rq, st) : launch3d_t2<S, k3Star, SO>(rq, st);
}
#endif

}  // namespace ssam

#define SSAM_INSTANTIATE_3D_PAIR(S, SO)                                                        \
    template int ssam::dispatch3d_pair<S, SO>(const ssam::Req3D&, hipStream_t);                \
    template int ssam::dispatch3d_t2_pair<S, SO>(const ssam::Req3D&, hipStream_t);
