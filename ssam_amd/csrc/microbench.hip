// microbench.hip -- CDNA4 micro-benchmarks behind the kernels' design choices:
// latency of one dependent operation on one wave (cycles) and device throughput
// (operations per cycle per CU) for the FP32 FMA, the DPP lane shift fused into
// an add (the systolic hop), ds_bpermute (LDS-crossbar lane permute), an LDS
// store + load round trip, and the 16x16x32 BF16 MFMA.  The reference measures
// the CUDA counterparts (shuffle vs shared memory vs FMA) for its performance
// model; these give the same evidence for wave64.
//
// Latency: one wave runs a chain of N dependent operations between two reads of
// the shader cycle counter (s_memtime).  Throughput: every CU runs 8 waves of 8
// independent chains; the time comes from s_memtime around the loop of each
// wave (max over waves), so no host clock conversion is needed.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "ssam_device.hpp"
#include "ssam_internal.hpp"

namespace ssam {

enum MicroOp : int {
    kMbFmaLat = 0,
    kMbDppLat = 1,
    kMbBpermLat = 2,
    kMbLdsLat = 3,
    kMbFmaTput = 4,
    kMbDppTput = 5,
    kMbBpermTput = 6,
    kMbLdsTput = 7,
    kMbMfmaTput = 8,
    // device rates in TFLOP/s from HIP events around a long kernel (wall clock, no
    // cycle counter): FP32 FMA (v_fmac_f32, no packed FP32), FP64 FMA, and the FP32 /
    // FP64 matrix instructions the banded-GEMM forms use
    kMbFmaF32Rate = 9,
    kMbFmaF64Rate = 10,
    kMbMfmaF32x4Rate = 11,   // v_mfma_f32_16x16x4_f32
    kMbMfmaF32x1Rate = 12,   // v_mfma_f32_4x4x1_16b_f32
    kMbMfmaF64Rate = 13,     // v_mfma_f64_16x16x4_f64
    // FP32 FMA operand forms of the stencil inner loop (TFLOP/s, wall clock):
    kMbFmacVgprRate = 14,    // v_fmac_f32 acc, v_tap, v_x (every operand a VGPR)
    kMbFmacSgprRate = 15,    // v_fmac_f32 acc, s_tap, v_x (one uniform tap)
    kMbFmacTapsRate = 16,    // v_fmac_f32 acc, s_tap[k], v_x over 16 kernel-argument taps
    kMbFmacDppRate = 17,     // 4 FMAs per DPP-fused add (the systolic hop's share)
    // the hop's issue cost among double-rate FMAs (TFLOP/s of the FMAs only): groups
    // of 10 VGPR-operand FMAs plus 2 DPP-fused adds, or plus 2 plain adds
    kMbFmacVgprDppRate = 18,
    kMbFmacVgprAddRate = 19,
    kMbCount = 20
};

constexpr int kMbIters = 1024;  // dependent operations per chain
constexpr int kMbChains = 8;    // independent chains per lane (throughput)

__device__ __forceinline__ uint64_t cycles() { return __builtin_amdgcn_s_memtime(); }

template <int OP>
__device__ __forceinline__ float chain_op(float x, float a, int lane, float* sh) {
    if constexpr (OP == kMbFmaLat || OP == kMbFmaTput) {
        return __builtin_fmaf(x, a, 1.0f);
    } else if constexpr (OP == kMbDppLat || OP == kMbDppTput) {
        return from_left(x) + a;
    } else if constexpr (OP == kMbBpermLat || OP == kMbBpermTput) {
        return __int_as_float(__builtin_amdgcn_ds_bpermute(((lane + 1) & 63) << 2, __float_as_int(x))) + a;
    } else {
        sh[lane] = x;
        __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): the store has reached LDS
        return sh[(lane + 1) & 63] + a;
    }
}

// out[0] = elapsed cycles of the chain, out[1] = sink (keeps the chain live)
template <int OP>
__global__ __launch_bounds__(64) void mb_latency_kernel(float a, unsigned long long* out, float* sink) {
    __shared__ float sh[64];
    const int lane = threadIdx.x;
    float x = static_cast<float>(lane) * 1e-3f;
    // a per-lane (VGPR) operand, so the DPP move can fuse into the add
    a *= 1.0f + static_cast<float>(lane) * 1e-7f;
    // warm the instruction cache and LDS with one short chain
#pragma unroll 4
    for (int i = 0; i < 32; ++i) x = chain_op<OP>(x, a, lane, sh);
    const uint64_t t0 = cycles();
#pragma unroll 16
    for (int i = 0; i < kMbIters; ++i) x = chain_op<OP>(x, a, lane, sh);
    const uint64_t t1 = cycles();
    if (lane == 0) out[0] = t1 - t0;
    sink[lane] = x;
}

typedef __bf16 mb_bf16x8 __attribute__((ext_vector_type(8)));
typedef float mb_f32x4 __attribute__((ext_vector_type(4)));

// out[block * 4 + wave] = cycles of that wave's loop
template <int OP>
__global__ __launch_bounds__(256) void mb_tput_kernel(float a, unsigned long long* out, float* sink) {
    __shared__ float sh[4][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    a *= 1.0f + static_cast<float>(lane) * 1e-7f;
    float x[kMbChains];
    mb_f32x4 acc[kMbChains];
#pragma unroll
    for (int c = 0; c < kMbChains; ++c) {
        x[c] = static_cast<float>(lane + c) * 1e-3f;
        acc[c] = mb_f32x4{0.f, 0.f, 0.f, 0.f};
    }
    mb_bf16x8 va, vb;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        va[e] = static_cast<__bf16>(a);
        vb[e] = static_cast<__bf16>(1.0f);
    }
    const uint64_t t0 = cycles();
    for (int i = 0; i < kMbIters / 8; ++i) {
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int c = 0; c < kMbChains; ++c) {
                if constexpr (OP == kMbMfmaTput)
                    acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va, vb, acc[c], 0, 0, 0);
                else
                    x[c] = chain_op<OP>(x[c], a, lane, sh[wv]);
            }
    }
    const uint64_t t1 = cycles();
    float s = 0.f;
#pragma unroll
    for (int c = 0; c < kMbChains; ++c) s += x[c] + acc[c][0] + acc[c][3];
    if (lane == 0) out[blockIdx.x * 4 + wv] = t1 - t0;
    sink[blockIdx.x * 256 + threadIdx.x] = s;
}

typedef double mb_f64x4 __attribute__((ext_vector_type(4)));
constexpr int kMbRateIters = 4096;  // loop trips of the rate kernels (8 chains each)

// Each lane runs 8 independent chains; `flops` per chain step is fixed by OP.
template <int OP>
__global__ __launch_bounds__(256) void mb_rate_kernel(float a, float* sink) {
    const int lane = threadIdx.x;
    const float fa = a * (1.0f + static_cast<float>(lane) * 1e-7f);
    float xf[kMbChains];
    double xd[kMbChains];
    mb_f32x4 af[kMbChains];
    mb_f64x4 ad[kMbChains];
#pragma unroll
    for (int c = 0; c < kMbChains; ++c) {
        xf[c] = static_cast<float>(lane + c) * 1e-3f;
        xd[c] = static_cast<double>(xf[c]);
        af[c] = mb_f32x4{0.f, 0.f, 0.f, 0.f};
        ad[c] = mb_f64x4{0.0, 0.0, 0.0, 0.0};
    }
    const double da = static_cast<double>(fa);
    for (int i = 0; i < kMbRateIters; ++i) {
#pragma unroll
        for (int c = 0; c < kMbChains; ++c) {
            if constexpr (OP == kMbFmaF32Rate) xf[c] = __builtin_fmaf(xf[c], fa, 1.0f);
            else if constexpr (OP == kMbFmaF64Rate) xd[c] = __builtin_fma(xd[c], da, 1.0);
            else if constexpr (OP == kMbMfmaF32x4Rate)
                af[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa, xf[c], af[c], 0, 0, 0);
            else if constexpr (OP == kMbMfmaF32x1Rate)
                af[c] = __builtin_amdgcn_mfma_f32_4x4x1f32(fa, xf[c], af[c], 0, 0, 0);
            else
                ad[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(da, xd[c], ad[c], 0, 0, 0);
        }
    }
    float s = 0.f;
#pragma unroll
    for (int c = 0; c < kMbChains; ++c)
        s += xf[c] + static_cast<float>(xd[c]) + af[c][0] + af[c][3] + static_cast<float>(ad[c][0] + ad[c][3]);
    sink[blockIdx.x * 256 + threadIdx.x] = s;
}

// The stencil's FMA forms: 8 accumulators per lane, data x[c] fixed per lane; the
// taps come from VGPRs, one SGPR, or 16 kernel-argument taps read as SGPRs.
struct MbTaps { float t[16]; };
template <int OP>
__global__ __launch_bounds__(256) void mb_fmac_kernel(MbTaps tp, float* sink) {
    const int lane = threadIdx.x;
    float x[kMbChains], acc[kMbChains];
    const float tv = tp.t[0] * (1.0f + static_cast<float>(lane) * 1e-7f);  // lane-varying tap
#pragma unroll
    for (int c = 0; c < kMbChains; ++c) {
        x[c] = static_cast<float>(lane + c) * 1e-3f;
        acc[c] = 0.f;
    }
    for (int i = 0; i < kMbRateIters; ++i) {
#pragma unroll
        for (int c = 0; c < kMbChains; ++c) {
            if constexpr (OP == kMbFmacVgprRate) acc[c] = __builtin_fmaf(tv, x[c], acc[c]);
            else if constexpr (OP == kMbFmacSgprRate) acc[c] = __builtin_fmaf(tp.t[0], x[c], acc[c]);
            else if constexpr (OP == kMbFmacTapsRate) acc[c] = __builtin_fmaf(tp.t[(c + 8 * (i & 1)) & 15], x[c], acc[c]);
            else {
                acc[c] = __builtin_fmaf(tp.t[c], x[c], acc[c]);
                if ((c & 3) == 3) acc[c] = from_left(acc[c - 3]) + acc[c];
            }
        }
    }
    float s = 0.f;
#pragma unroll
    for (int c = 0; c < kMbChains; ++c) s += acc[c];
    sink[blockIdx.x * 256 + threadIdx.x] = s;
}

// 10 accumulators per lane with VGPR taps; after each group of 10 FMAs, two
// partials cross one lane (DPP add, the systolic hop) or are added in place.
constexpr int kMbHopGroup = 10;
template <int OP>
__global__ __launch_bounds__(256) void mb_hop_kernel(float a, float* sink) {
    const int lane = threadIdx.x & (kWave - 1);
    float x[kMbHopGroup], acc[kMbHopGroup], tv[kMbHopGroup];
#pragma unroll
    for (int c = 0; c < kMbHopGroup; ++c) {
        x[c] = static_cast<float>(lane + c) * 1e-3f;
        acc[c] = 0.f;
        tv[c] = a * (1.0f + static_cast<float>(c) * 1e-3f + static_cast<float>(lane) * 1e-7f);
    }
    for (int i = 0; i < kMbRateIters; ++i) {
#pragma unroll
        for (int c = 0; c < kMbHopGroup; ++c) acc[c] = __builtin_fmaf(tv[c], x[c], acc[c]);
        if constexpr (OP == kMbFmacVgprDppRate) {
            acc[0] += from_left(acc[4]);
            acc[9] += from_right(acc[5]);
        } else {
            acc[0] += acc[4];
            acc[9] += acc[5];
        }
    }
    float s = 0.f;
#pragma unroll
    for (int c = 0; c < kMbHopGroup; ++c) s += acc[c];
    sink[blockIdx.x * 256 + threadIdx.x] = s;
}

// flops of one chain step of one wave
template <int OP> constexpr double mb_wave_flops() {
    return OP == kMbFmaF32Rate || OP == kMbFmaF64Rate || OP >= kMbFmacVgprRate ? 2.0 * 64
           : OP == kMbMfmaF32x1Rate                   ? 2.0 * 16 * 4 * 4 * 1
                                                      : 2.0 * 16 * 16 * 4;
}

}  // namespace ssam

using namespace ssam;

namespace {

template <int OP> int run_latency(double* result, hipStream_t st) {
    unsigned long long* d_out = nullptr;
    float* d_sink = nullptr;
    if (hipMalloc(&d_out, sizeof(unsigned long long)) != hipSuccess) return SSAM_E_HIP;
    if (hipMalloc(&d_sink, 64 * sizeof(float)) != hipSuccess) {
        (void)hipFree(d_out);
        return SSAM_E_HIP;
    }
    unsigned long long cyc = 0;
    int rc = SSAM_OK;
    for (int rep = 0; rep < 2 && rc == SSAM_OK; ++rep) {  // second run: warm caches
        hipLaunchKernelGGL(mb_latency_kernel<OP>, dim3(1), dim3(64), 0, st, 1.0001f, d_out, d_sink);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(&cyc, d_out, sizeof(cyc), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            rc = SSAM_E_HIP;
    }
    (void)hipFree(d_out);
    (void)hipFree(d_sink);
    *result = static_cast<double>(cyc) / kMbIters;
    return rc;
}

template <int OP> int run_tput(double* result, hipStream_t st) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        return SSAM_E_HIP;
    const int blocks = 2 * cus;  // 8 waves per CU
    unsigned long long* d_out = nullptr;
    float* d_sink = nullptr;
    if (hipMalloc(&d_out, blocks * 4 * sizeof(unsigned long long)) != hipSuccess) return SSAM_E_HIP;
    if (hipMalloc(&d_sink, blocks * 256 * sizeof(float)) != hipSuccess) {
        (void)hipFree(d_out);
        return SSAM_E_HIP;
    }
    std::vector<unsigned long long> cyc(blocks * 4);
    int rc = SSAM_OK;
    for (int rep = 0; rep < 2 && rc == SSAM_OK; ++rep) {
        hipLaunchKernelGGL(mb_tput_kernel<OP>, dim3(blocks), dim3(256), 0, st, 1.0001f, d_out, d_sink);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(cyc.data(), d_out, cyc.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            rc = SSAM_E_HIP;
    }
    (void)hipFree(d_out);
    (void)hipFree(d_sink);
    unsigned long long mx = 1;
    for (auto c : cyc) mx = c > mx ? c : mx;
    // wave-instructions issued per CU (8 waves x iterations x chains) over the
    // slowest wave's cycles
    *result = 8.0 * kMbIters * kMbChains / static_cast<double>(mx);
    return rc;
}

template <int OP> int run_rate(double* result, hipStream_t st) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        return SSAM_E_HIP;
    const int blocks = 8 * cus;  // 32 waves per CU
    float* d_sink = nullptr;
    if (hipMalloc(&d_sink, static_cast<size_t>(blocks) * 256 * sizeof(float)) != hipSuccess) return SSAM_E_HIP;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = SSAM_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = SSAM_E_HIP;
    float best = 1e30f;
    for (int rep = 0; rep < 3 && rc == SSAM_OK; ++rep) {  // first run warms the clocks
        float ms = 0.f;
        if (hipEventRecord(e0, st) != hipSuccess) { rc = SSAM_E_HIP; break; }
        if constexpr (OP == kMbFmacVgprDppRate || OP == kMbFmacVgprAddRate) {
            hipLaunchKernelGGL(mb_hop_kernel<OP>, dim3(blocks), dim3(256), 0, st, 1.0001f, d_sink);
        } else if constexpr (OP >= kMbFmacVgprRate) {
            MbTaps tp;
            for (int k = 0; k < 16; ++k) tp.t[k] = 1.0001f + 1e-4f * k;
            hipLaunchKernelGGL(mb_fmac_kernel<OP>, dim3(blocks), dim3(256), 0, st, tp, d_sink);
        } else {
            hipLaunchKernelGGL(mb_rate_kernel<OP>, dim3(blocks), dim3(256), 0, st, 1.0001f, d_sink);
        }
        if (hipGetLastError() != hipSuccess || hipEventRecord(e1, st) != hipSuccess ||
            hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            rc = SSAM_E_HIP;
        else if (rep > 0 && ms < best)
            best = ms;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(d_sink);
    const int per_iter = OP == kMbFmacVgprDppRate || OP == kMbFmacVgprAddRate ? kMbHopGroup : kMbChains;
    const double flops = static_cast<double>(blocks) * 4 * kMbRateIters * per_iter * mb_wave_flops<OP>();
    *result = rc == SSAM_OK ? flops / (static_cast<double>(best) * 1e-3) / 1e12 : 0.0;
    return rc;
}

}  // namespace

extern "C" int ssam_microbench(int which, double* result, void* stream) {
    if (!result || which < 0 || which >= kMbCount) return SSAM_E_INVALID;
    hipStream_t st = static_cast<hipStream_t>(stream);
    switch (which) {
        case kMbFmaLat: return run_latency<kMbFmaLat>(result, st);
        case kMbDppLat: return run_latency<kMbDppLat>(result, st);
        case kMbBpermLat: return run_latency<kMbBpermLat>(result, st);
        case kMbLdsLat: return run_latency<kMbLdsLat>(result, st);
        case kMbFmaTput: return run_tput<kMbFmaTput>(result, st);
        case kMbDppTput: return run_tput<kMbDppTput>(result, st);
        case kMbBpermTput: return run_tput<kMbBpermTput>(result, st);
        case kMbLdsTput: return run_tput<kMbLdsTput>(result, st);
        case kMbMfmaTput: return run_tput<kMbMfmaTput>(result, st);
        case kMbFmaF32Rate: return run_rate<kMbFmaF32Rate>(result, st);
        case kMbFmaF64Rate: return run_rate<kMbFmaF64Rate>(result, st);
        case kMbMfmaF32x4Rate: return run_rate<kMbMfmaF32x4Rate>(result, st);
        case kMbMfmaF32x1Rate: return run_rate<kMbMfmaF32x1Rate>(result, st);
        case kMbMfmaF64Rate: return run_rate<kMbMfmaF64Rate>(result, st);
        case kMbFmacVgprRate: return run_rate<kMbFmacVgprRate>(result, st);
        case kMbFmacSgprRate: return run_rate<kMbFmacSgprRate>(result, st);
        case kMbFmacTapsRate: return run_rate<kMbFmacTapsRate>(result, st);
        case kMbFmacDppRate: return run_rate<kMbFmacDppRate>(result, st);
        case kMbFmacVgprDppRate: return run_rate<kMbFmacVgprDppRate>(result, st);
        default: return run_rate<kMbFmacVgprAddRate>(result, st);
    }
}
