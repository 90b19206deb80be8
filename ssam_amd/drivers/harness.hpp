// harness.hpp -- host side shared by the three command-line drivers
// (ssai-2dstencil, ssai-3dstencil, ssai-2dconv), which keep the reference
// harnesses' command lines and output lines and call the kernels through the C
// ABI (include/ssam.h).
//
//  * Image<T>: host grid with the reference's raw file formats -- 8-bit raw image
//    load (ssai-2dconv/common.h:148-178, `../data/Lena<W>x<H>.raw`), raw dump
//    (common.h:111-123) and the nested-brace text dump (common.h:124-147);
//  * the harness self-check: a CPU correlation with zero padding and the 3D
//    interior-only sweep, run after the GPU call like the reference harness does
//    (stencil2D_5pt-v1.cu:215-228, tmp3d7pt-v2.cu:312-320);
//  * seeded inputs (splitmix64) and HIP helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ssam.h"

namespace harness {

#define HIP_OK(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            std::fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, \
                         __LINE__);                                                       \
            std::exit(3);                                                                 \
        }                                                                                 \
    } while (0)

#define SSAM_OK_OR_DIE(expr)                                                              \
    do {                                                                                  \
        int s_ = (expr);                                                                  \
        if (s_ != SSAM_OK) {                                                              \
            std::fprintf(stderr, "ssam error: %s (%d) at %s:%d\n", ssam_strerror(s_), s_,    \
                         __FILE__, __LINE__);                                             \
            std::exit(4);                                                                 \
        }                                                                                 \
    } while (0)

inline uint64_t splitmix64(uint64_t seed, uint64_t i) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
// U[0,1) with 24 random bits
inline double u01(uint64_t seed, uint64_t i) {
    return static_cast<double>(splitmix64(seed, i) >> 40) * (1.0 / 16777216.0);
}

template <typename T> struct Image {
    int width = 0, height = 0;
    std::vector<T> data;
    void alloc(int w, int h) {
        width = w;
        height = h;
        data.assign(static_cast<size_t>(w) * h, T(0));
    }
    // 8-bit raw image (common.h:148-178); false if the file is absent or short
    bool load_uchar(const char* path, int w, int h) {
        FILE* fp = std::fopen(path, "rb");
        if (!fp) return false;
        std::vector<unsigned char> b(static_cast<size_t>(w) * h);
        const size_t got = std::fread(b.data(), 1, b.size(), fp);
        std::fclose(fp);
        if (got != b.size()) return false;
        alloc(w, h);
        for (size_t i = 0; i < b.size(); ++i) data[i] = static_cast<T>(b[i]);
        return true;
    }
    bool save_raw(const char* path) const {
        FILE* fp = std::fopen(path, "wb");
        bool ok = false;
        if (fp) {
            ok = std::fwrite(data.data(), sizeof(T), data.size(), fp) == data.size();
            std::fclose(fp);
        }
        std::printf("save:%s,%s\n", path, ok ? "success" : "failed");
        return ok;
    }
    void save_text(const char* path) const {
        std::printf("save : %s\n", path);
        FILE* fp = std::fopen(path, "wt");
        if (!fp) return;
        std::fprintf(fp, "{\n {\n  {\n");
        size_t s = 0;
        for (int m = 0; m < height; ++m) {
            std::fprintf(fp, "   {");
            for (int n = 0; n < width; ++n) std::fprintf(fp, "%f ", static_cast<float>(data[s++]));
            std::fprintf(fp, "   }\n");
        }
        std::fprintf(fp, "   }\n }\n}\n");
        std::fclose(fp);
    }
};

// Harness self-check: zero-padded correlation, filter column outer, row inner,
// accumulated in T -- the loop the reference harness checks against.
template <typename T>
void correlate(const T* src, T* dst, int width, int height, const T* filter, int fw, int fh) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < height; ++i)
        for (int j = 0; j < width; ++j) {
            T sum = 0;
            for (int n = 0; n < fw; ++n)
                for (int m = 0; m < fh; ++m) {
                    const int x = j + n - fw / 2, y = i + m - fh / 2;
                    if (x >= 0 && x < width && y >= 0 && y < height)
                        sum += src[static_cast<size_t>(y) * width + x] * filter[m * fw + n];
                }
            dst[static_cast<size_t>(i) * width + j] = sum;
        }
}

template <typename T> struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    explicit DevBuf(size_t count) : n(count) { HIP_OK(hipMalloc(&p, count * sizeof(T))); }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    void upload(const T* h) { HIP_OK(hipMemcpy(p, h, n * sizeof(T), hipMemcpyHostToDevice)); }
    void download(T* h) const { HIP_OK(hipMemcpy(h, p, n * sizeof(T), hipMemcpyDeviceToHost)); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

struct EventTimer {
    hipEvent_t a, b;
    EventTimer() {
        HIP_OK(hipEventCreate(&a));
        HIP_OK(hipEventCreate(&b));
    }
    ~EventTimer() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
    void start() { HIP_OK(hipEventRecord(a, nullptr)); }
    float stop_ms() {
        HIP_OK(hipEventRecord(b, nullptr));
        HIP_OK(hipEventSynchronize(b));
        float ms = 0;
        HIP_OK(hipEventElapsedTime(&ms, a, b));
        return ms;
    }
};

}  // namespace harness
