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
    kMbCount = 9
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
        default: return run_tput<kMbMfmaTput>(result, st);
    }
}
