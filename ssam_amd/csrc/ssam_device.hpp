// ssam_device.hpp -- device-side building blocks shared by the SSAM kernels
// (gfx950, wave64).
//
//  * storage <-> accumulation conversion for F32 / F64 / BF16 / F16 grids;
//  * whole-wave DPP lane shifts (wave_shr:1 / wave_shl:1), the systolic hop that
//    replaces the reference's 32-lane __shfl_up/down_sync
//    (ssai-2dconv/cudaLib.cuh:573-581);
//  * V-wide per-lane row strip loads/stores (one dwordx4 per lane when the strip
//    is 16 bytes and aligned), with the zero / clamp boundary rules.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace ssam {

constexpr int kWave = 64;

// ---------------------------------------------------------------- storage types
struct bf16_t { uint16_t bits; };
struct f16_t { uint16_t bits; };

template <typename S> struct Storage;
template <> struct Storage<float> {
    using acc = float;
    __device__ static inline float load(float v) { return v; }
    __device__ static inline float store(float v) { return v; }
};
template <> struct Storage<double> {
    using acc = double;
    __device__ static inline double load(double v) { return v; }
    __device__ static inline double store(double v) { return v; }
};
template <> struct Storage<bf16_t> {
    using acc = float;
    __device__ static inline float load(bf16_t v) {
        return __uint_as_float(static_cast<uint32_t>(v.bits) << 16);
    }
    // round to nearest even (v_cvt_pk_bf16_f32 on gfx950)
    __device__ static inline bf16_t store(float v) {
        __bf16 b = static_cast<__bf16>(v);
        bf16_t r;
        __builtin_memcpy(&r.bits, &b, 2);
        return r;
    }
};
template <> struct Storage<f16_t> {
    using acc = float;
    __device__ static inline float load(f16_t v) {
        _Float16 h;
        __builtin_memcpy(&h, &v.bits, 2);
        return static_cast<float>(h);
    }
    __device__ static inline f16_t store(float v) {
        // the FP32 result is rounded to FP16 as a value of its own: without the
        // opaque copy the compiler folds the last FMA and the conversion into
        // v_fma_mixlo_f16 (one rounding of a*b+c straight to FP16), which can
        // differ from rounding the FP32 sum in a tie, so the same sweeps would
        // round differently in a fused launch and in a chain of launches
        asm("" : "+v"(v));
        _Float16 h = static_cast<_Float16>(v);
        f16_t r;
        __builtin_memcpy(&r.bits, &h, 2);
        return r;
    }
};

// ---------------------------------------------------------------- fused multiply-add
// in the accumulation type (v_fmac_f32 / v_fma_f64); never promotes FP32 to FP64
__device__ __forceinline__ float fmaa(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaa(double a, double b, double c) { return __builtin_fma(a, b, c); }

// ---------------------------------------------------------------- DPP lane shifts
// DPP controls (GFX9 encoding): wave_shl:1 = 0x130 (lane i reads lane i+1),
// wave_shr:1 = 0x138 (lane i reads lane i-1).  Lanes shifted in from outside the
// wave read 0 (bound_ctrl); callers only consume them on halo lanes.
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
    const uint64_t u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(u & 0xffffffffu), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(u >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double((static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) |
                                static_cast<uint32_t>(lo));
}
// value of lane i+1 (data moves toward lower lanes)
template <typename A> __device__ __forceinline__ A from_right(A v) { return dpp_move<0x130>(v); }
// value of lane i-1 (data moves toward higher lanes)
template <typename A> __device__ __forceinline__ A from_left(A v) { return dpp_move<0x138>(v); }

// ---------------------------------------------------------------- global access
// Every grid access goes through a global (address space 1) pointer.  A pointer
// read from memory (a batch table) is otherwise generic and becomes a flat
// access, which also counts against the scalar-memory counter: each scalar load
// would then wait for every load in flight.
template <typename T> struct RawBits { using type = T; };
template <> struct RawBits<bf16_t> { using type = uint16_t; };
template <> struct RawBits<f16_t> { using type = uint16_t; };

template <typename T> __device__ __forceinline__ T gld(const void* p) {
    using R = typename RawBits<T>::type;
    const R r = *(const __attribute__((address_space(1))) R*)p;
    T v;
    __builtin_memcpy(&v, &r, sizeof(T));
    return v;
}
template <typename T> __device__ __forceinline__ void gst(void* p, T v) {
    using R = typename RawBits<T>::type;
    R r;
    __builtin_memcpy(&r, &v, sizeof(T));
    *(__attribute__((address_space(1))) R*)p = r;
}

// ---------------------------------------------------------------- raw buffers
// Buffer resource over [base, base + bytes) (bytes < 2^31).  An access whose
// per-lane byte offset is at or past `bytes` reads 0 and drops its store (the
// hardware range check of raw buffers, offset = VGPR offset + immediate), so an
// offset built from an out-of-grid coordinate, which wraps to >= 2^31, never
// touches memory outside the buffer.  The SGPR offset operand is always 0.
using BufRsrc = __amdgpu_buffer_rsrc_t;
constexpr uint32_t kBufOOB = 0x80000000u;  // an offset no buffer here reaches

__device__ __forceinline__ BufRsrc buf_rsrc(const void* base, uint32_t bytes) {
    // dword 3: 32-bit data format, the value for gfx9-family raw access
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), static_cast<short>(0),
                                             static_cast<int>(bytes), 0x00020000);
}

template <int N> struct BufWords;
template <> struct BufWords<1> { using type = uint32_t; };
template <> struct BufWords<2> { typedef uint32_t type __attribute__((ext_vector_type(2))); };
template <> struct BufWords<4> { typedef uint32_t type __attribute__((ext_vector_type(4))); };

// N consecutive 4-byte elements (or N/2 doubles) at byte offset `off`; 8 dwords
// go as two 4-dword accesses
template <typename T, int N>
__device__ __forceinline__ void buf_load(BufRsrc r, uint32_t off, T (&out)[N]) {
    constexpr int W = N * static_cast<int>(sizeof(T)) / 4;
    if constexpr (W == 8) {
        T (&lo)[N / 2] = *reinterpret_cast<T (*)[N / 2]>(&out[0]);
        T (&hi)[N / 2] = *reinterpret_cast<T (*)[N / 2]>(&out[N / 2]);
        buf_load<T, N / 2>(r, off, lo);
        buf_load<T, N / 2>(r, off + 16u, hi);
    } else {
        static_assert(W == 1 || W == 2 || W == 4, "1, 2, 4 or 8 dwords");
        typename BufWords<W>::type v;
        if constexpr (W == 1) v = __builtin_amdgcn_raw_buffer_load_b32(r, static_cast<int>(off), 0, 0);
        else if constexpr (W == 2) v = __builtin_amdgcn_raw_buffer_load_b64(r, static_cast<int>(off), 0, 0);
        else v = __builtin_amdgcn_raw_buffer_load_b128(r, static_cast<int>(off), 0, 0);
        __builtin_memcpy(&out[0], &v, sizeof(v));
    }
}

template <typename T, int N>
__device__ __forceinline__ void buf_store(BufRsrc r, uint32_t off, const T (&in)[N]) {
    constexpr int W = N * static_cast<int>(sizeof(T)) / 4;
    if constexpr (W == 8) {
        const T (&lo)[N / 2] = *reinterpret_cast<const T (*)[N / 2]>(&in[0]);
        const T (&hi)[N / 2] = *reinterpret_cast<const T (*)[N / 2]>(&in[N / 2]);
        buf_store<T, N / 2>(r, off, lo);
        buf_store<T, N / 2>(r, off + 16u, hi);
    } else {
        static_assert(W == 1 || W == 2 || W == 4, "1, 2, 4 or 8 dwords");
        typename BufWords<W>::type v;
        __builtin_memcpy(&v, &in[0], sizeof(v));
        if constexpr (W == 1) __builtin_amdgcn_raw_buffer_store_b32(v, r, static_cast<int>(off), 0, 0);
        else if constexpr (W == 2) __builtin_amdgcn_raw_buffer_store_b64(v, r, static_cast<int>(off), 0, 0);
        else __builtin_amdgcn_raw_buffer_store_b128(v, r, static_cast<int>(off), 0, 0);
    }
}

// V consecutive elements of storage type S at byte offset `off` of a raw buffer,
// in the accumulation type.  VEC: one access for the V elements (their bytes a
// multiple of 4); otherwise an access per element (16-bit ones for BF16 / F16).
template <typename S, int V, bool VEC>
__device__ __forceinline__ void buf_load_elems(BufRsrc r, uint32_t off, typename Storage<S>::acc (&out)[V]) {
    using A = typename Storage<S>::acc;
    constexpr uint32_t ES = sizeof(S);
    if constexpr (sizeof(S) == sizeof(A)) {
        if constexpr (VEC) {
            buf_load<A, V>(r, off, out);
        } else {
#pragma unroll
            for (int j = 0; j < V; ++j) {
                A e[1];
                buf_load<A, 1>(r, off + static_cast<uint32_t>(j) * ES, e);
                out[j] = e[0];
            }
        }
    } else if constexpr (VEC && (V * sizeof(S)) % 4 == 0) {
        uint32_t w[V / 2];
        buf_load<uint32_t, V / 2>(r, off, w);
#pragma unroll
        for (int j = 0; j < V; ++j) {
            S s;
            s.bits = static_cast<uint16_t>((j & 1) ? (w[j / 2] >> 16) : (w[j / 2] & 0xffffu));
            out[j] = Storage<S>::load(s);
        }
    } else {
#pragma unroll
        for (int j = 0; j < V; ++j) {
            S s;
            s.bits = __builtin_amdgcn_raw_buffer_load_b16(r, static_cast<int>(off + static_cast<uint32_t>(j) * ES), 0, 0);
            out[j] = Storage<S>::load(s);
        }
    }
}

// Stores V values (accumulation type) as storage type S: VEC, one access at
// so[0] + roff; otherwise element j at so[j] + roff (an offset at or past the
// buffer end drops that element).
template <typename S, int V, bool VEC>
__device__ __forceinline__ void buf_store_elems(BufRsrc r, const uint32_t (&so)[V], uint32_t roff,
                                                const typename Storage<S>::acc (&v)[V]) {
    using A = typename Storage<S>::acc;
    if constexpr (sizeof(S) == sizeof(A)) {
        if constexpr (VEC) {
            buf_store<A, V>(r, so[0] + roff, v);
        } else {
#pragma unroll
            for (int j = 0; j < V; ++j) {
                const A e[1] = {v[j]};
                buf_store<A, 1>(r, so[j] + roff, e);
            }
        }
    } else if constexpr (VEC && (V * sizeof(S)) % 4 == 0) {
        uint32_t w[V / 2];
#pragma unroll
        for (int j = 0; j < V; j += 2)
            w[j / 2] = static_cast<uint32_t>(Storage<S>::store(v[j]).bits) |
                       (static_cast<uint32_t>(Storage<S>::store(v[j + 1]).bits) << 16);
        buf_store<uint32_t, V / 2>(r, so[0] + roff, w);
    } else {
#pragma unroll
        for (int j = 0; j < V; ++j)
            __builtin_amdgcn_raw_buffer_store_b16(Storage<S>::store(v[j]).bits, r, static_cast<int>(so[j] + roff), 0, 0);
    }
}

// ---------------------------------------------------------------- strip I/O
// One lane's V consecutive elements of a row, as one vector access.  The vector
// type only promises the element's alignment: rows of any pitch take the same
// 8/16-byte access (gfx9-family global memory runs unaligned accesses, in the
// unaligned mode the ROCm runtime sets), so only the element alignment of the
// base pointer is required.
template <typename S, int V> struct VecBits {
    static constexpr int bytes = V * static_cast<int>(sizeof(S));
    static_assert(bytes % 4 == 0, "strip must be a whole number of dwords");
    typedef uint32_t vec_t __attribute__((ext_vector_type(bytes / 4)));
    typedef vec_t type __attribute__((aligned(sizeof(S))));
};

template <typename S, int V>
__device__ __forceinline__ void load_strip_vec(const S* __restrict__ p, typename Storage<S>::acc (&out)[V]) {
    using VB = typename VecBits<S, V>::type;
    VB raw = gld<VB>(p);
    S tmp[V];
    __builtin_memcpy(tmp, &raw, sizeof(tmp));
#pragma unroll
    for (int j = 0; j < V; ++j) out[j] = Storage<S>::load(tmp[j]);
}

template <typename S, int V>
__device__ __forceinline__ void store_strip_vec(S* __restrict__ p, const typename Storage<S>::acc (&v)[V]) {
    using VB = typename VecBits<S, V>::type;
    S tmp[V];
#pragma unroll
    for (int j = 0; j < V; ++j) tmp[j] = Storage<S>::store(v[j]);
    VB raw;
    __builtin_memcpy(&raw, tmp, sizeof(tmp));
    gst<VB>(p, raw);
}

// One lane's V elements of a row as the 2D input ring holds them: F32 / F64 as
// values; BF16 / F16 packed two to a dword as loaded, converted to FP32 where a
// tick reads them.  A ring slot loaded in one ring period and first read in the
// next then crosses the loop back-edge as the load left it, so the memory-counter
// wait sits at the read (exact) instead of at the back-edge, where converting
// would wait for every row still in flight once per period.
template <typename S, int V> struct RowReg {
    using A = typename Storage<S>::acc;
    static constexpr bool kPacked = sizeof(S) == 2;
    static constexpr int N = kPacked ? V / 2 : V;
    using E = typename std::conditional<kPacked, uint32_t, A>::type;
    static_assert(!kPacked || V % 2 == 0, "packed rows hold element pairs");
    E r[N];

    __device__ __forceinline__ void zero() {
#pragma unroll
        for (int n = 0; n < N; ++n) r[n] = E(0);
    }
    // one vector access (element-aligned p)
    __device__ __forceinline__ void load_vec(const S* p) {
        if constexpr (kPacked) {
            using VB = typename VecBits<S, V>::type;
            const VB raw = gld<VB>(p);
            __builtin_memcpy(r, &raw, sizeof(r));
        } else {
            load_strip_vec<S, V>(p, r);
        }
    }
    // one raw-buffer access of the V elements at byte offset `off`: every dword
    // at or past the buffer's end reads 0 (the range check of raw buffers is per
    // dword), and so does an offset at or past 2^31
    __device__ __forceinline__ void load_buf(BufRsrc rs, uint32_t off) {
        if constexpr (kPacked) buf_load<uint32_t, N>(rs, off, r);
        else buf_load<A, V>(rs, off, r);
    }
    // from V elements in storage form (edge paths)
    __device__ __forceinline__ void set(const S (&e)[V]) {
        if constexpr (kPacked) {
#pragma unroll
            for (int n = 0; n < N; ++n) {
                uint16_t lo, hi;
                __builtin_memcpy(&lo, &e[2 * n], 2);
                __builtin_memcpy(&hi, &e[2 * n + 1], 2);
                r[n] = static_cast<uint32_t>(lo) | (static_cast<uint32_t>(hi) << 16);
            }
        } else {
#pragma unroll
            for (int j = 0; j < V; ++j) r[j] = Storage<S>::load(e[j]);
        }
    }
    __device__ __forceinline__ A get(int j) const {
        if constexpr (!kPacked) {
            return r[j];
        } else if constexpr (std::is_same<S, bf16_t>::value) {
            return __uint_as_float((j & 1) ? (r[j / 2] & 0xffff0000u) : (r[j / 2] << 16));
        } else {
            const uint16_t b = static_cast<uint16_t>((j & 1) ? (r[j / 2] >> 16) : (r[j / 2] & 0xffffu));
            _Float16 h;
            __builtin_memcpy(&h, &b, 2);
            return static_cast<float>(h);
        }
    }
};

// Boundary rules for reads outside the grid.
enum Boundary : int { kZero = 0, kClamp = 1, kHold = 2, kDirichlet0 = 3 };

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

}  // namespace ssam
