// util.hip -- seeded synthetic input and the stream-copy bandwidth kernel.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "ssam_device.hpp"
#include "ssam_internal.hpp"

namespace ssam {

// splitmix64 (Steele, Lea, Flood 2014): element i is the (i+1)-th output of a
// stream whose state starts at `seed`, so any element is computed directly.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

template <typename S>
__global__ __launch_bounds__(256) void fill_uniform_kernel(S* __restrict__ dst, int64_t n,
                                                           uint64_t seed, uint64_t offset,
                                                           double a, double b) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += stride) {
        const uint64_t u = mix64(seed + (offset + static_cast<uint64_t>(i) + 1) * 0x9E3779B97F4A7C15ULL);
        const double v = a + (b - a) * (static_cast<double>(u >> 40) * (1.0 / 16777216.0));
        if constexpr (sizeof(S) == 8) {
            dst[i] = v;
        } else if constexpr (std::is_same<S, float>::value) {
            dst[i] = static_cast<float>(v);
        } else {
            dst[i] = Storage<S>::store(static_cast<float>(v));
        }
    }
}

// dwordx4 grid-stride copy with non-temporal loads and stores (streamed data
// bypasses the caches' retention): the measured HBM stream bandwidth.  Four
// independent 16-byte loads per thread are in flight before the stores.
typedef uint32_t v4u __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void stream_copy_kernel(const v4u* __restrict__ src,
                                                          v4u* __restrict__ dst, int64_t n16) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n16; i += 4 * stride) {
        const v4u v0 = __builtin_nontemporal_load(src + i);
        const v4u v1 = __builtin_nontemporal_load(src + i + stride);
        const v4u v2 = __builtin_nontemporal_load(src + i + 2 * stride);
        const v4u v3 = __builtin_nontemporal_load(src + i + 3 * stride);
        __builtin_nontemporal_store(v0, dst + i);
        __builtin_nontemporal_store(v1, dst + i + stride);
        __builtin_nontemporal_store(v2, dst + i + 2 * stride);
        __builtin_nontemporal_store(v3, dst + i + 3 * stride);
    }
    for (; i < n16; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

}  // namespace ssam

extern "C" int ssam_fill_uniform(void* dst, int dtype, int64_t n, uint64_t seed, uint64_t offset,
                                 double a, double b, void* stream) {
    using namespace ssam;
    if (!dst || n < 0) return SSAM_E_INVALID;
    if (n == 0) return SSAM_OK;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int blocks = static_cast<int>(n / 256 + 1 < 8192 ? n / 256 + 1 : 8192);
    switch (dtype) {
        case SSAM_F32:
            hipLaunchKernelGGL(fill_uniform_kernel<float>, dim3(blocks), dim3(256), 0, st,
                               static_cast<float*>(dst), n, seed, offset, a, b);
            break;
        case SSAM_F64:
            hipLaunchKernelGGL(fill_uniform_kernel<double>, dim3(blocks), dim3(256), 0, st,
                               static_cast<double*>(dst), n, seed, offset, a, b);
            break;
        case SSAM_BF16:
            hipLaunchKernelGGL(fill_uniform_kernel<bf16_t>, dim3(blocks), dim3(256), 0, st,
                               static_cast<bf16_t*>(dst), n, seed, offset, a, b);
            break;
        case SSAM_F16:
            hipLaunchKernelGGL(fill_uniform_kernel<f16_t>, dim3(blocks), dim3(256), 0, st,
                               static_cast<f16_t*>(dst), n, seed, offset, a, b);
            break;
        default:
            return SSAM_E_INVALID;
    }
    return hipGetLastError() == hipSuccess ? SSAM_OK : SSAM_E_HIP;
}

extern "C" int ssam_stream_copy(const void* src, void* dst, int64_t bytes, void* stream) {
    using namespace ssam;
    if (!src || !dst || bytes < 0 || bytes % 16 != 0) return SSAM_E_INVALID;
    if ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) % 16 != 0)
        return SSAM_E_INVALID;
    if (bytes == 0) return SSAM_OK;
    const int64_t n16 = bytes / 16;
    // up to 64 workgroups per CU (256 CUs); each thread moves 4 x 16 B per pass
    const int64_t want = (n16 + 255) / 256;
    const int blocks = static_cast<int>(want < 16384 ? want : 16384);
    hipLaunchKernelGGL(stream_copy_kernel, dim3(blocks), dim3(256), 0,
                       static_cast<hipStream_t>(stream), static_cast<const v4u*>(src),
                       static_cast<v4u*>(dst), n16);
    return hipGetLastError() == hipSuccess ? SSAM_OK : SSAM_E_HIP;
}
