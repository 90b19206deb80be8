// stencil2d_mfma.hpp -- large 2D boxes and filters on the matrix cores (gfx950
// v_mfma_f32_16x16x32_{bf16,f16}) for BF16 / FP16 storage, FP32 accumulation.
//
// Banded-GEMM form of the correlation out(y,x) = sum_{m,n} w[m][n] in(y+m-r, x+n-r):
// for a 16x16 output tile at (y0, x0) and filter row m,
//     D(i,j) += sum_k A_m(i,k) B_m(k,j),   A_m(i,k) = in(y0+i+m-r, x0-8+k),
//     B_m(k,j) = w[m][k-8+r-j]  (0 <= k-8+r-j <= 2r, else 0),  k = 0..31,
// so one MFMA per filter row covers the 16 + 2r <= 32 input columns of the tile
// (r <= 8; r = 9 takes two 32-column windows).  The input window starts 8 (16)
// columns left of the tile, so every lane's A operand (8 consecutive columns of
// one row) is one aligned 16-byte load.
//
// Register cache instead of LDS: lane (i, g) of a 16-lane row group holds, for
// each of the wave's NT tiles side by side, input rows as 16-row blocks in a
// ring; A_m for row shift m is built from blocks b and b+1 with two DPP moves
// (row_shl:m from block b, row_shr:16-m from block b+1) -- the systolic hop of
// the VALU kernels, run down the rows.  The wave walks down a segment of rows in
// steps of 16; each step loads one new 16-row block per tile, NB-2 steps ahead.
//
// Taps: FP32 from the kernel-argument segment, converted to the storage type;
// FP16 storage splits them into hi + lo parts (two MFMAs per filter row) so the
// products carry ~22 mantissa bits of each tap; BF16 storage rounds its result
// to 8 bits anyway and uses one part (SSAM_MFMA_SPLIT=1 splits it too); taps
// exact in the storage type never split.  Inputs are exact in the storage type,
// accumulation is FP32.
//
// Output D(i, j) sits in lane j + 16*(i/4), element i%4 (16x16 result layout);
// rows at or past the segment end and columns past the width are not stored.
// Zero boundary: rows and columns outside the grid read as 0.
#pragma once
#include <cstdlib>

#include "ssam_device.hpp"
#include "ssam_internal.hpp"
#include "stencil2d.hpp"

namespace ssam {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
// 8 consecutive 2-byte elements of a row of any pitch (element alignment only;
// the access is one unaligned 16-byte load)
typedef u32x4 u32x4_row __attribute__((aligned(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <typename S> struct MfmaOp;
template <> struct MfmaOp<bf16_t> {
    typedef __bf16 v8 __attribute__((ext_vector_type(8)));
    __device__ static __forceinline__ f32x4 mma(u32x4 a, u32x4 b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8, a), __builtin_bit_cast(v8, b),
                                                       c, 0, 0, 0);
    }
};
template <> struct MfmaOp<f16_t> {
    typedef _Float16 v8 __attribute__((ext_vector_type(8)));
    __device__ static __forceinline__ f32x4 mma(u32x4 a, u32x4 b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(v8, a), __builtin_bit_cast(v8, b),
                                                      c, 0, 0, 0);
    }
};

struct ParamsM {
    int width, height;
    long long pitch;  // elements between rows
    int nstrips;      // strips of 16*NT output columns
    int seg_rows;     // output rows per tile (multiple of 16)
    int ntiles, nblocks;
};

constexpr int kWavesPerBlockM = 4;

// Rows shifted up by M inside each 16-lane row group: lane i gets x from lane
// i+M while i+M < 16, else y from lane i+M-16 (x = rows b, y = rows b+1).
template <int M> __device__ __forceinline__ uint32_t shift_rows(uint32_t x, uint32_t y) {
    if constexpr (M == 0) {
        return x;
    } else if constexpr (M == 16) {
        return y;
    } else {
        // both moves fill lanes whose source leaves the row group with 0, so the
        // two halves combine with an OR (foldable into v_or_b32 with a DPP source)
        const uint32_t lo = static_cast<uint32_t>(
            __builtin_amdgcn_update_dpp(0, static_cast<int>(x), 0x100 + M, 0xf, 0xf, true));
        const uint32_t hi = static_cast<uint32_t>(
            __builtin_amdgcn_update_dpp(0, static_cast<int>(y), 0x110 + (16 - M), 0xf, 0xf, true));
        return lo | hi;
    }
}

// NC 32-column windows per 16-column tile: 1 for r <= 8, 2 for r = 9 (the window
// then starts 16 columns left of the tile).
template <typename S, int R, bool SPLIT, int NT, int NC>
struct Mfma2D {
    static_assert(R >= 1 && 16 + 2 * R <= 32 * NC && R <= 8 * NC, "window must hold 16 + 2r columns");
    static_assert(2 * R <= 32, "rows s..s+2 of the ring hold the 16 + 2r rows of a step");
    static constexpr int K = 2 * R + 1;
    static constexpr int NB = 4;        // row blocks in the ring (two used + two ahead)
    static constexpr int OW = 16 * NT;  // output columns per strip
    static constexpr int XS = 8 * NC;   // window start, columns left of the tile
    static constexpr int NA = NT * NC;  // A operands per row block
    using Op = MfmaOp<S>;

    struct Lane {
        const char* src;
        char* dst;
        ParamsM p;
        int i, g;      // A layout: row i of the block, column chunk g
        int x0;        // first output column of the strip
        int ys, ye;    // output rows [ys, ye) of the tile
        int blast;     // last row block any output row needs
        uint32_t gbytes;  // bytes from the grid's first element to past its last
    };

    // Row block b (rows ys - R + 16 b + i) of operand a = t*NC + c (tile t, window
    // c): 8 columns from x0 + 16t - XS + 32c + 8g.
    // Edge tiles (CHK): one raw-buffer access over the whole grid per chunk, with
    // no branch (per-lane branches around element loads made the compiler wait
    // for every load in flight).  Rows outside the grid and chunks left of it or
    // past the right edge take an offset past the buffer and read 0; a chunk
    // straddling the right edge reads on into the row's padding or the next row
    // (inside the buffer) and its columns past the edge are cleared.  The dword
    // holding the grid's very last element (odd width) also holds the element
    // past the buffer's end, so the range check drops it: that element comes on
    // its own.  The grid spans < 2 GiB (mfma_eligible).
    template <bool CHK>
    __device__ static __forceinline__ u32x4 load_block(const Lane& L, int b, int a) {
        const int r = L.ys - R + 16 * min(b, L.blast) + L.i;
        const int c = L.x0 + 16 * (a / NC) - XS + 32 * (a % NC) + 8 * L.g;
        if constexpr (!CHK) {
            const char* row = L.src + static_cast<long long>(r) * L.p.pitch * static_cast<long long>(sizeof(S));
            return gld<u32x4_row>(row + static_cast<long long>(c) * static_cast<long long>(sizeof(S)));
        } else {
            const bool ok = r >= 0 && r < L.p.height && c >= 0 && c < L.p.width;
            const uint32_t off = ok ? (static_cast<uint32_t>(r) * static_cast<uint32_t>(L.p.pitch) +
                                       static_cast<uint32_t>(c)) * 2u
                                    : kBufOOB;
            const BufRsrc rs = buf_rsrc(L.src, L.gbytes);
            uint32_t w[4];
            buf_load<uint32_t, 4>(rs, off, w);
            const int k = L.p.width - 1 - c;  // the row's last column within this chunk
            const bool last_of_grid = ok && r == L.p.height - 1 && (L.p.width & 1) && k < 8;
            const uint32_t lastv = __builtin_amdgcn_raw_buffer_load_b16(
                rs, static_cast<int>(last_of_grid ? off + static_cast<uint32_t>(k) * 2u : kBufOOB), 0, 0);
            u32x4 v;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                // columns c+2q (low half) and c+2q+1 (high half) past the edge read 0
                const uint32_t lo = 2 * q <= k ? (w[q] & 0xffffu) : 0u;
                const uint32_t hi = 2 * q + 1 <= k ? (w[q] & 0xffff0000u) : 0u;
                v[q] = last_of_grid && k == 2 * q ? lastv : (lo | hi);
            }
            return v;
        }
    }

    // B operand of filter row m, window c (hi or lo part), built once per wave.
    __device__ static __forceinline__ void build_taps(const float* __restrict__ w, int j, int g,
                                                      u32x4 (&bh)[K * NC], u32x4 (&bl)[K * NC]) {
#pragma unroll
        for (int mc = 0; mc < K * NC; ++mc) {
            const int m = mc / NC, c = mc % NC;
            uint32_t h[8], l[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int n = 32 * c + 8 * g + e - XS + R - j;
                const float v = (n >= 0 && n <= 2 * R) ? w[m * K + n] : 0.f;
                const S hi = Storage<S>::store(v);
                const S lo = Storage<S>::store(v - Storage<S>::load(hi));
                h[e] = hi.bits;
                l[e] = lo.bits;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                bh[mc][q] = h[2 * q] | (h[2 * q + 1] << 16);
                bl[mc][q] = l[2 * q] | (l[2 * q + 1] << 16);
            }
        }
    }

    template <int M>
    __device__ static __forceinline__ u32x4 shifted(const u32x4& x, const u32x4& y) {
        u32x4 a;
#pragma unroll
        for (int q = 0; q < 4; ++q) a[q] = shift_rows<M>(x[q], y[q]);
        return a;
    }

    // hi and lo parts accumulate separately (independent MFMA chains)
    // X, Y, Z: row blocks s, s+1, s+2 (Z is read only when 2r > 16)
    template <int M>
    __device__ static __forceinline__ void rows_from(f32x4 (&D)[NT], f32x4 (&E)[NT], const u32x4 (&X)[NA],
                                                     const u32x4 (&Y)[NA], const u32x4 (&Z)[NA],
                                                     const u32x4 (&bh)[K * NC], const u32x4 (&bl)[K * NC]) {
        if constexpr (M < K) {
#pragma unroll
            for (int t = 0; t < NT; ++t)
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    const int q = t * NC + c;
                    const u32x4 a = M <= 16 ? shifted<(M <= 16 ? M : 0)>(X[q], Y[q])
                                            : shifted<(M > 16 ? M - 16 : 0)>(Y[q], Z[q]);
                    D[t] = Op::mma(a, bh[M * NC + c], D[t]);
                    if constexpr (SPLIT) E[t] = Op::mma(a, bl[M * NC + c], E[t]);
                }
            rows_from<M + 1>(D, E, X, Y, Z, bh, bl);
        }
    }

    template <bool CHK>
    __device__ static __forceinline__ void store_tile(const Lane& L, int y0, int t, const f32x4& d) {
        const int j = L.i;            // result layout: column = lane % 16
        const int rg = L.g;           // rows 4*rg .. 4*rg+3
        const int x = L.x0 + 16 * t + j;
        // interior tiles: ye - ys is a whole number of 16-row steps, so every row
        // is stored unconditionally (straight-line code keeps the memory counter
        // waits exact, so prefetched row blocks stay in flight); edge tiles store
        // through a raw buffer over the grid, outputs outside the tile taking an
        // offset past it (no branch either)
        [[maybe_unused]] BufRsrc rs;
        if constexpr (CHK) rs = buf_rsrc(L.dst, L.gbytes);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int y = y0 + 4 * rg + q;
            if constexpr (CHK) {
                const bool ok = y < L.ye && x < L.p.width;
                const uint32_t off = ok ? (static_cast<uint32_t>(y) * static_cast<uint32_t>(L.p.pitch) +
                                           static_cast<uint32_t>(x)) * 2u
                                        : kBufOOB;
                __builtin_amdgcn_raw_buffer_store_b16(Storage<S>::store(d[q]).bits, rs, static_cast<int>(off), 0, 0);
            } else {
                char* p = L.dst + (static_cast<long long>(y) * L.p.pitch + x) * static_cast<long long>(sizeof(S));
                gst<S>(p, Storage<S>::store(d[q]));
            }
        }
    }

    // One 16-row step s (ring slot KT = s mod NB): prefetch block s+NB-1 into the
    // slot of block s-1, then the MFMAs over blocks s and s+1, then the stores.
    template <bool CHK, int KT>
    __device__ static __forceinline__ void step(const Lane& L, int s, u32x4 (&ring)[NB][NA],
                                                const u32x4 (&bh)[K * NC], const u32x4 (&bl)[K * NC]) {
#pragma unroll
        for (int a = 0; a < NA; ++a) ring[(KT + NB - 1) % NB][a] = load_block<CHK>(L, s + NB - 1, a);
        f32x4 D[NT], E[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) D[t] = E[t] = f32x4{0.f, 0.f, 0.f, 0.f};
        rows_from<0>(D, E, ring[KT % NB], ring[(KT + 1) % NB], ring[(KT + 2) % NB], bh, bl);
        const int y0 = L.ys + 16 * s;
#pragma unroll
        for (int t = 0; t < NT; ++t) store_tile<CHK>(L, y0, t, SPLIT ? D[t] + E[t] : D[t]);
    }

    // a full ring period: NB steps, straight-line code
    template <bool CHK, int KT>
    __device__ static __forceinline__ void period(const Lane& L, int s0, u32x4 (&ring)[NB][NA],
                                                  const u32x4 (&bh)[K * NC], const u32x4 (&bl)[K * NC]) {
        if constexpr (KT < NB) {
            step<CHK, KT>(L, s0 + KT, ring, bh, bl);
            period<CHK, KT + 1>(L, s0, ring, bh, bl);
        }
    }

    // the last partial period: steps s0+KT < nsteps only
    template <bool CHK, int KT>
    __device__ static __forceinline__ void tail(const Lane& L, int s0, int nsteps, u32x4 (&ring)[NB][NA],
                                                const u32x4 (&bh)[K * NC], const u32x4 (&bl)[K * NC]) {
        if constexpr (KT < NB - 1) {
            if (s0 + KT < nsteps) {
                step<CHK, KT>(L, s0 + KT, ring, bh, bl);
                tail<CHK, KT + 1>(L, s0, nsteps, ring, bh, bl);
            }
        }
    }

    template <bool CHK>
    __device__ static __forceinline__ void stream(const Lane& L, const u32x4 (&bh)[K * NC],
                                                  const u32x4 (&bl)[K * NC]) {
        u32x4 ring[NB][NA];
#pragma unroll
        for (int b = 0; b < NB - 1; ++b)
#pragma unroll
            for (int a = 0; a < NA; ++a) ring[b][a] = load_block<CHK>(L, b, a);
        const int nsteps = (L.ye - L.ys + 15) / 16;
        int s0 = 0;
        for (; s0 + NB <= nsteps; s0 += NB) period<CHK, 0>(L, s0, ring, bh, bl);
        tail<CHK, 0>(L, s0, nsteps, ring, bh, bl);
    }

    // One tile: strip `strip` (OW output columns), output rows [seg*seg_rows, ...)
    // of the grid described by p (wave-uniform arguments).
    __device__ static __forceinline__ void run_tile(const S* __restrict__ src, S* __restrict__ dst,
                                                    const ParamsM& p, const float* __restrict__ w,
                                                    int strip, int seg) {
        const int lane = threadIdx.x & (kWave - 1);
        Lane L;
        L.src = reinterpret_cast<const char*>(src);
        L.dst = reinterpret_cast<char*>(dst);
        L.p = p;
        L.i = lane & 15;
        L.g = lane >> 4;
        L.x0 = strip * OW;
        L.ys = seg * p.seg_rows;
        L.ye = min(p.height, L.ys + p.seg_rows);
        // output rows up to ye-1 need input rows up to ye-1+R: block index of that row
        L.blast = (L.ye - 1 + R - (L.ys - R)) / 16;
        L.gbytes = static_cast<uint32_t>(((static_cast<long long>(p.height) - 1) * p.pitch + p.width) * 2);
        u32x4 bh[K * NC], bl[K * NC];
        build_taps(w, L.i, L.g, bh, bl);
        const bool interior = L.x0 - XS >= 0 && L.x0 + OW - 16 - XS + 32 * NC <= p.width && L.ys - R >= 0 &&
                              L.ys - R + 16 * L.blast + 15 < p.height && (L.ye - L.ys) % 16 == 0;
        if (interior)
            stream<false>(L, bh, bl);
        else
            stream<true>(L, bh, bl);
    }

    __device__ static void run(const S* __restrict__ src, S* __restrict__ dst, const ParamsM& p,
                               const float* __restrict__ w) {
        const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
        const int tile = xcd_remap(blockIdx.x, p.nblocks) * kWavesPerBlockM + wv;
        if (tile >= p.ntiles) return;  // whole wave leaves; no barriers are used
        run_tile(src, dst, p, w, tile % p.nstrips, tile / p.nstrips);
    }
};

// one set of taps in one window leaves room for 3 waves per SIMD, split taps for
// 2; two windows (radius 9) need the whole register file (1)
template <typename S, int R, bool SPLIT, int NT, int NC>
__global__ __launch_bounds__(kWave * kWavesPerBlockM)
__attribute__((amdgpu_waves_per_eu(NC > 1 ? 1 : (SPLIT ? 2 : 3))))
void mfma2d_kernel(Taps<float, (2 * R + 1) * (2 * R + 1)> tp, const S* __restrict__ src,
                   S* __restrict__ dst, ParamsM p) {
    (void)tp;
    const float* __restrict__ w = (const float*)(__builtin_amdgcn_kernarg_segment_ptr());
    Mfma2D<S, R, SPLIT, NT, NC>::run(src, dst, p, w);
}

// Host launch for one request; returns SSAM_E_UNSUPPORTED when the request is
// outside what the matrix-core form covers.
template <typename S, int R, bool SPLIT>
int launch_mfma_instance(const Req2D& rq, hipStream_t st) {
    constexpr int NC = R <= 8 ? 1 : 2;  // 32-column windows per tile
    constexpr int NT = 3 - NC;          // tiles side by side per wave (register budget)
    constexpr int K = 2 * R + 1;
    Taps<float, K * K> tp;
    for (int i = 0; i < K * K; ++i) tp.w[i] = static_cast<float>(rq.taps[i]);
    ParamsM p;
    p.width = static_cast<int>(rq.width);
    p.height = static_cast<int>(rq.height);
    p.pitch = rq.pitch;
    p.nstrips = static_cast<int>((rq.width + 16 * NT - 1) / (16 * NT));
    int seg = rq.seg_rows;
    if (seg <= 0) {
        const long long target_waves = static_cast<long long>(rq.waves_per_simd) * 1024;
        long long nsegs = (target_waves + p.nstrips - 1) / p.nstrips;
        if (nsegs < 1) nsegs = 1;
        seg = static_cast<int>((rq.height + nsegs - 1) / nsegs);
        if (seg < 64) seg = 64;
    }
    // whole ring periods (4 steps of 16 rows) keep interior tiles on the
    // straight-line loop
    seg = (seg + 63) / 64 * 64;
    if (seg > (p.height + 15) / 16 * 16) seg = (p.height + 15) / 16 * 16;
    p.seg_rows = seg;
    const long long ntiles = static_cast<long long>(p.nstrips) * ((p.height + seg - 1) / seg);
    if (ntiles > (1LL << 30)) return SSAM_E_UNSUPPORTED;
    p.ntiles = static_cast<int>(ntiles);
    p.nblocks = (p.ntiles + kWavesPerBlockM - 1) / kWavesPerBlockM;
    hipLaunchKernelGGL((mfma2d_kernel<S, R, SPLIT, NT, NC>), dim3(p.nblocks), dim3(kWave * kWavesPerBlockM), 0,
                       st, tp, static_cast<const S*>(rq.src), static_cast<S*>(rq.dst), p);
    return hipGetLastError() == hipSuccess ? SSAM_OK : SSAM_E_HIP;
}

// Whether every tap is exact in the storage type (then one MFMA per row).
template <typename S> bool taps_exact(const Req2D& rq) {
    const int K = 2 * rq.radius + 1;
    for (int i = 0; i < K * K; ++i) {
        const float v = static_cast<float>(rq.taps[i]);
        if (static_cast<double>(v) != rq.taps[i]) return false;
        uint32_t u;
        __builtin_memcpy(&u, &v, 4);
        if constexpr (std::is_same<S, bf16_t>::value) {
            if (u & 0xFFFFu) return false;  // more than 8 mantissa bits
        } else {
            const _Float16 h = static_cast<_Float16>(v);
            if (static_cast<float>(h) != v) return false;
        }
    }
    return true;
}

// Matrix-core form: BF16/FP16 storage, box/dense radius 1..9, one sweep, zero
// boundary; rows of any pitch.
template <typename S> bool mfma_eligible(const Req2D& rq) {
    if constexpr (!std::is_same<S, bf16_t>::value && !std::is_same<S, f16_t>::value) {
        (void)rq;
        return false;
    } else {
        // radius 9 needs two windows per tile and the whole register file, so its
        // taps are never split: FP16 taps rounded to 11 bits add at most 2^-11 of
        // sum |w x| to the 2^-11 of the FP16 output rounding (inside 1e-3)
        const uintptr_t a = reinterpret_cast<uintptr_t>(rq.src) | reinterpret_cast<uintptr_t>(rq.dst);
        // (edge tiles address the grid through one raw buffer: under 2 GiB)
        const long long bytes = ((rq.height - 1) * rq.pitch + rq.width) * 2;
        return rq.shape == kBox && rq.steps == 1 && rq.boundary == kZero && rq.radius >= 1 &&
               rq.radius <= 9 && a % 2 == 0 && bytes < (1LL << 31);
    }
}

// Taps split hi + lo unless they are exact in the storage type.  BF16 storage
// rounds its result to 8 mantissa bits, so BF16 taps (another 2^-9 relative at
// most) keep it well inside its error budget; SSAM_MFMA_SPLIT=1 forces the split.
template <typename S> bool mfma_split(const Req2D& rq) {
    if (taps_exact<S>(rq) || rq.radius > 8) return false;
    if (std::is_same<S, bf16_t>::value) {
        const char* v = std::getenv("SSAM_MFMA_SPLIT");
        return v && *v == '1';
    }
    return true;
}

template <typename S> int launch_mfma(const Req2D& rq, hipStream_t st) {
    if constexpr (!std::is_same<S, bf16_t>::value && !std::is_same<S, f16_t>::value) {
        (void)rq;
        (void)st;
        return SSAM_E_UNSUPPORTED;
    } else {
        const bool split = mfma_split<S>(rq);
        switch (rq.radius) {
#define SSAM_RM(r)                                                                              \
    case r:                                                                                     \
        return split ? launch_mfma_instance<S, r, true>(rq, st) : launch_mfma_instance<S, r, false>(rq, st);
            SSAM_RM(1) SSAM_RM(2) SSAM_RM(3) SSAM_RM(4) SSAM_RM(5) SSAM_RM(6) SSAM_RM(7) SSAM_RM(8)
#undef SSAM_RM
            case 9:
                return launch_mfma_instance<S, 9, false>(rq, st);
            default:
                return SSAM_E_UNSUPPORTED;
        }
    }
}

}  // namespace ssam
