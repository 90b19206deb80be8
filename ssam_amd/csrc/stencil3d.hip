// stencil3d.hip -- launch and dispatch of the 3D kernels (stencil3d.hpp).
#include <hip/hip_runtime.h>

#include <cstdlib>

#include "ssam_internal.hpp"
#include "stencil3d.hpp"
#include "stencil3d_hl.hpp"

namespace ssam {

// fixed-order sum of the per-tile partials (one workgroup); out[0] = the sum, or
// its square root (the L2 norm) when take_sqrt is set
__global__ __launch_bounds__(256) void sum_partials_kernel(const double* __restrict__ part, int n,
                                                           double* __restrict__ out, int take_sqrt) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += part[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = take_sqrt ? sqrt(sh[0]) : sh[0];
}

template <typename S> struct Lanes3D;
template <> struct Lanes3D<float> { static constexpr int V = 4; };
template <> struct Lanes3D<double> { static constexpr int V = 2; };

template <typename S, int SHAPE, int R, int V>
static int launch3d(const Req3D& rq, hipStream_t st) {
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
               (reinterpret_cast<uintptr_t>(rq.src) % 16 == 0) &&
               (reinterpret_cast<uintptr_t>(rq.dst) % 16 == 0);
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
    hipLaunchKernelGGL((ssam3d_kernel<S, SHAPE, R, V, NY, PF>), dim3(p.nblocks),
                       dim3(kWave * kWaves3D), 0, st, tp, static_cast<const S*>(rq.src),
                       static_cast<S*>(rq.dst), p);
    if (hipGetLastError() != hipSuccess) return SSAM_E_HIP;
    if (rq.partials) {
        hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, st, rq.partials,
                           p.ntiles, rq.res_out, rq.res_sqrt);
        if (hipGetLastError() != hipSuccess) return SSAM_E_HIP;
    }
    return SSAM_OK;
}

template <typename S>
static int dispatch3d_t(const Req3D& rq, hipStream_t st) {
    constexpr int V = Lanes3D<S>::V;
    if (rq.shape == k3FaceEdge) {
        if (rq.radius != 1) return SSAM_E_UNSUPPORTED;
        return launch3d<S, k3FaceEdge, 1, V>(rq, st);
    }
    if (rq.radius == 1) return launch3d<S, k3Star, 1, V>(rq, st);
    if (rq.radius == 2) return launch3d<S, k3Star, 2, V>(rq, st);
    return SSAM_E_UNSUPPORTED;
}

// ---------------------------------------------------------------- two sweeps per launch
// Temporal blocking for radius-1 stars and the 19-point face+edge stencil on
// grids under 2 GiB: one launch runs two sweeps (Stencil3DHL2) with one HBM read
// and one write of the grid.  Tiles (x-strip, y-block, z-segment) in that order,
// one per wave; blocks remapped so each XCD works on a contiguous range of
// tiles (neighbouring y-blocks share their halo rows in that XCD's L2).
template <typename S, int SHAPE, int V, int NY, int PF>
__global__ __launch_bounds__(kWave * kWaves3D) __attribute__((amdgpu_waves_per_eu(2)))
void ssam3d_t2_kernel(Taps3<typename Storage<S>::acc, ntaps3(SHAPE, 1)> tp, const S* __restrict__ src,
                      S* __restrict__ dst, Grid3D g, int nstrips, int nyb, int zseg, int ntiles, int nblocks,
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
    Stencil3DHL2<S, SHAPE, V, NY, PF>::run_tile(src, dst, g, vec != 0, w, strip, yb, zs, ze, false);
}

template <typename S, int SHAPE>
static int launch3d_t2(const Req3D& rq, hipStream_t st) {
    using A = typename Storage<S>::acc;
    // FP32: two columns per lane; FP64: one (two halo lanes per side)
    constexpr int V = sizeof(S) == 8 ? 1 : 2;
    constexpr int NY = 6, PF = 1;
    using K = Stencil3DHL2<S, SHAPE, V, NY, PF>;
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
    const uintptr_t al = static_cast<uintptr_t>(V) * sizeof(S);
    const int vec = (rq.nx % V == 0) && (reinterpret_cast<uintptr_t>(rq.src) % al == 0) &&
                    (reinterpret_cast<uintptr_t>(rq.dst) % al == 0);
    hipLaunchKernelGGL((ssam3d_t2_kernel<S, SHAPE, V, NY, PF>), dim3(nblocks), dim3(kWave * kWaves3D), 0, st, tp,
                       static_cast<const S*>(rq.src), static_cast<S*>(rq.dst), g, nstrips, nyb, zseg,
                       static_cast<int>(ntiles), nblocks, vec);
    return hipGetLastError() == hipSuccess ? SSAM_OK : SSAM_E_HIP;
}

bool two_sweeps_ok(int storage, const Req3D& rq) {
    return (storage == SSAM_F32 || storage == SSAM_F64) && rq.radius == 1 &&
           grid3d_fits_buffer(rq.nx, rq.ny, rq.nz, storage == SSAM_F64 ? 8 : 4);
}

int dispatch3d_two_sweeps(int storage, const Req3D& rq, hipStream_t st) {
    if (!two_sweeps_ok(storage, rq)) return SSAM_E_UNSUPPORTED;
    if (storage == SSAM_F32)
        return rq.shape == k3FaceEdge ? launch3d_t2<float, k3FaceEdge>(rq, st) : launch3d_t2<float, k3Star>(rq, st);
    return rq.shape == k3FaceEdge ? launch3d_t2<double, k3FaceEdge>(rq, st) : launch3d_t2<double, k3Star>(rq, st);
}

int dispatch3d(int storage, const Req3D& rq, hipStream_t st) {
    if (storage == SSAM_F32) return dispatch3d_t<float>(rq, st);
    if (storage == SSAM_F64) return dispatch3d_t<double>(rq, st);
    return SSAM_E_UNSUPPORTED;
}

int partials_needed(int storage, const Req3D& rq) {
    const int V = storage == SSAM_F64 ? 2 : 4;
    const int NY = rq.radius == 1 ? 4 : 2;
    const long long nstrips = (rq.nx + kWave * V - 1) / (kWave * V);
    const long long nyb = (rq.ny + NY - 1) / NY;
    // worst case: the heuristic segment length (at least 4r planes)
    const long long xy = nstrips * nyb;
    long long nzs = (4LL * 1024 + xy - 1) / xy;
    long long zseg = (rq.nz + nzs - 1) / nzs;
    if (zseg < 4 * rq.radius) zseg = 4 * rq.radius;
    if (rq.zseg > 0) zseg = rq.zseg;
    if (zseg > rq.nz) zseg = rq.nz;
    return static_cast<int>(nstrips * nyb * ((rq.nz + zseg - 1) / zseg));
}

}  // namespace ssam
