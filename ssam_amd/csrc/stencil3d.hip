// stencil3d.hip -- dispatch of the 3D kernels (stencil3d.hpp, stencil3d_hl.hpp)
// by the storage pair of a launch, and the residual's fixed-order sum.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include <algorithm>

// (the kernel instances of every storage pair are in their own units:
// stencil3d_f32f64.hip, stencil3d_bf16.hip, stencil3d_f16.hip)
#include "stencil3d_launch.hpp"

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

int launch_sum_partials(const double* part, int n, double* out, int take_sqrt, hipStream_t st) {
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, st, part, n, out, take_sqrt);
    return hipGetLastError() == hipSuccess ? SSAM_OK : SSAM_E_HIP;
}

namespace {

// (src, dst) storage pairs: the same type, or a 2-byte type to / from the FP32
// running state of its chain
bool pair_ok(int in, int out) {
    if (in == out) return in == SSAM_F32 || in == SSAM_F64 || in == SSAM_BF16 || in == SSAM_F16;
    return (in == SSAM_F32 && (out == SSAM_BF16 || out == SSAM_F16)) ||
           (out == SSAM_F32 && (in == SSAM_BF16 || in == SSAM_F16));
}

int elem_bytes3(int storage) { return storage == SSAM_F64 ? 8 : storage == SSAM_F32 ? 4 : 2; }

}  // namespace

bool two_sweeps_ok(int storage, const Req3D& rq) {
    const int eb = std::max(elem_bytes3(storage), elem_bytes3(rq.out_storage));
    return pair_ok(storage, rq.out_storage) && rq.radius == 1 && grid3d_fits_buffer(rq.nx, rq.ny, rq.nz, eb);
}

int dispatch3d_two_sweeps(int storage, const Req3D& rq, hipStream_t st) {
    if (!two_sweeps_ok(storage, rq)) return SSAM_E_UNSUPPORTED;
    switch (storage * 4 + rq.out_storage) {
        case SSAM_F32 * 4 + SSAM_F32: return dispatch3d_t2_pair<float, float>(rq, st);
        case SSAM_F64 * 4 + SSAM_F64: return dispatch3d_t2_pair<double, double>(rq, st);
        case SSAM_BF16 * 4 + SSAM_BF16: return dispatch3d_t2_pair<bf16_t, bf16_t>(rq, st);
        case SSAM_BF16 * 4 + SSAM_F32: return dispatch3d_t2_pair<bf16_t, float>(rq, st);
        case SSAM_F32 * 4 + SSAM_BF16: return dispatch3d_t2_pair<float, bf16_t>(rq, st);
        case SSAM_F16 * 4 + SSAM_F16: return dispatch3d_t2_pair<f16_t, f16_t>(rq, st);
        case SSAM_F16 * 4 + SSAM_F32: return dispatch3d_t2_pair<f16_t, float>(rq, st);
        case SSAM_F32 * 4 + SSAM_F16: return dispatch3d_t2_pair<float, f16_t>(rq, st);
        default: return SSAM_E_UNSUPPORTED;
    }
}

int dispatch3d(int storage, const Req3D& rq, hipStream_t st) {
    if (!pair_ok(storage, rq.out_storage)) return SSAM_E_UNSUPPORTED;
    switch (storage * 4 + rq.out_storage) {
        case SSAM_F32 * 4 + SSAM_F32: return dispatch3d_pair<float, float>(rq, st);
        case SSAM_F64 * 4 + SSAM_F64: return dispatch3d_pair<double, double>(rq, st);
        case SSAM_BF16 * 4 + SSAM_BF16: return dispatch3d_pair<bf16_t, bf16_t>(rq, st);
        case SSAM_BF16 * 4 + SSAM_F32: return dispatch3d_pair<bf16_t, float>(rq, st);
        case SSAM_F32 * 4 + SSAM_BF16: return dispatch3d_pair<float, bf16_t>(rq, st);
        case SSAM_F16 * 4 + SSAM_F16: return dispatch3d_pair<f16_t, f16_t>(rq, st);
        case SSAM_F16 * 4 + SSAM_F32: return dispatch3d_pair<f16_t, float>(rq, st);
        case SSAM_F32 * 4 + SSAM_F16: return dispatch3d_pair<float, f16_t>(rq, st);
        default: return SSAM_E_UNSUPPORTED;
    }
}

int partials_needed(int storage, const Req3D& rq) {
    const int V = storage == SSAM_F64 ? 2 : 4;  // Lanes3D of the src storage
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
