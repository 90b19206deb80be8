// stencil2d.hpp -- SSAM register-cache systolic 2D stencil / correlation kernel
// for gfx950 (wave64).
//
// Replaces the reference kernels
//   stencil2d_v1::j2d5pt      (ssai-2dstencil/stencil2D_5pt-v1.cu:8-129)
//   stencil2d_{9,13,17,21,25}pt::j2d*pt (stencil2D_*pt.cu, star radius 2..6)
//   stencil2d_64pt/81/121pt::j2d*pt     (box radius 4 / 4 / 5)
//   stencil2d_tmp::j2d5pt / j2d9pt      (stencil2d_temp_{5,9}pt.cu, fused steps)
//   CONV2D::kernel_convolution2D        (ssai-2dconv/conv2D.cu:9-70)
// with one template.  All compute out(x,y) = sum_{n,m} w[n][m] in(x+m-r, y+n-r),
// i.e. correlation without flip, the same as the reference CPU gold
// (ssai-2dconv/common.h:186-216).
//
// Execution model (MI355X-first, not a port of the 32-lane code):
//  * One wave64 owns a column strip; lane l caches V consecutive columns
//    (one dwordx4 for F32 V=4, F64 V=2, BF16/F16 V=8).  HL = ceil(T*r/V) lanes on
//    each side are halo lanes: they load, compute and pass partial sums, but do
//    not store.  Useful fraction (64-2*HL)/64 (reference: (32-2r)/32).
//  * The wave streams down a segment of rows.  Each lane keeps a register ring of
//    2r+1 (+prefetch) input rows; every tick loads one new row (prefetched PF rows
//    ahead) and produces one output row per time step.  Each input element is
//    read from HBM once per segment (plus 2*T*r halo rows per segment).
//  * Horizontal taps: for output column j a lane forms the partial sum of its own
//    columns (d = 0) and, for neighbours at lane distance d, the partial that the
//    neighbour needs; partials travel lane to lane with DPP wave_shl:1 /
//    wave_shr:1, fused into v_add_f32_dpp (the systolic hop; no LDS, no shuffle
//    through memory).
//  * T time steps run as a T-stage pipeline of register rings in one launch
//    (temporal blocking); intermediate rows/columns outside the grid are forced to
//    zero each step, so the fused result equals T zero-padded sweeps exactly as in
//    the reference gold iteration (stencil2d_temp_5pt.cu:169-175).
//  * Workgroups are remapped so each XCD gets a contiguous band of tiles
//    (vertically adjacent segments share that XCD's L2 for their halo rows).
#pragma once
#include "ssam_device.hpp"
#include "ssam_internal.hpp"

namespace ssam {

template <typename A, int NT> struct Taps { A w[NT]; };

struct Params2D {
    int width, height;
    long long pitch;      // elements between rows of src
    long long pitch_out;  // elements between rows of dst (its own storage type)
    int nstrips;      // column strips
    int ntiles;       // nstrips * nsegs
    int seg_rows;     // output rows per tile
    int vec_ok;       // strips may use one vector access (pitch, pointers aligned)
    int nblocks;      // gridDim.x (for the XCD remap)
    int boundary;     // Boundary rule (kZero, kClamp, kHold)
    const void* taps;  // device taps (nullptr: the kernel-argument segment)
};

// The taps of one box as the kernel reads them: scalar loads at uniform,
// compile-time offsets from one base (the kernel-argument segment, a batch table
// or the caller's device array).  kBoxEven: an even square filter of 2r x 2r taps
// stored densely (pitch 2r) at the top-left of the (2r+1)^2 box, whose last row
// and column are zero and never read (the reference correlation's offset n - fw/2,
// ssai-2dconv/common.h:196-197).
template <typename A, int K, int SHAPE> struct TapView {
    using CP = const __attribute__((address_space(4))) A*;  // constant space: scalar loads
    CP w;
    static constexpr bool kRowPipe = false;  // taps loaded a filter row ahead
    static constexpr bool kVgpr = false;     // FMAs read the tap SGPRs
    static constexpr bool kSkip(int n, int c) { return SHAPE == kBoxEven && (n == K - 1 || c == K - 1); }
    __device__ __forceinline__ A at(int n, int c) const {
        return SHAPE == kBoxEven ? w[n * (K - 1) + c] : w[n * K + c];
    }
};

// Where an FP32 FMA takes its tap from.  On gfx950 v_fmac_f32 issues at twice
// the rate when every operand is a VGPR; an SGPR operand (a uniform tap) halves
// it (measured: 119 vs 73 TFLOP/s, ssam_microbench 14 / 15).  So FP32
// accumulation copies the taps into VGPRs: resident for the whole tile when the
// shape has at most kMaxResidentTaps of them, else once per tick (one v_mov per
// tap feeds its V FMAs) up to kMaxTickTaps; past that the register allocator
// keeps too many of the copies live and spills, so the largest filters and all
// FP64 FMAs read the tap SGPRs directly.  The per-tick VGPR form runs the
// filter one row at a time (stencil_row): row n+1's scalar loads are issued
// while row n computes, and a scheduling barrier between rows keeps the
// scheduler from hoisting every copy of a large filter at once (which spilled).
// The SGPR form keeps the scheduler's own order: barriers there exposed each
// row's load latency (FP64 box121 ran 2.3x slower).
enum TapMode : int { kTapSgpr = 0, kTapResident = 1, kTapTick = 2 };
constexpr int kMaxResidentTaps = 128;
constexpr int kMaxTickTaps = 196;
template <typename A, int R, int SHAPE> struct TapPolicy {
    static constexpr int K = 2 * R + 1;
    static constexpr int used = SHAPE == kStar ? 4 * R + 1 : SHAPE == kBoxEven ? 4 * R * R : K * K;
    // compensated boxes hold two more partial arrays; FP64 taps take two VGPRs each
    // and only small sets stay resident (no rate to gain, but no scalar loads)
    static constexpr int resident_max = sizeof(A) == 8 ? 32 : SHAPE == kBoxKahan ? 49 : kMaxResidentTaps;
    static constexpr int mode = used <= resident_max                     ? kTapResident
                                : sizeof(A) == 8 || used > kMaxTickTaps ? kTapSgpr
                                                                         : kTapTick;
};

// a uniform value as a VGPR operand (the asm output is opaque, so the FMAs that
// use it cannot fold the SGPR back in)
__device__ __forceinline__ float to_vgpr(float s) {
    float v;
    asm("v_mov_b32 %0, %1" : "=v"(v) : "s"(s));
    return v;
}
__device__ __forceinline__ double to_vgpr(double s) {
    double v;
    asm("v_mov_b64 %0, %1" : "=v"(v) : "s"(s));
    return v;
}

// per-tick VGPR copies of the taps (kTapTick)
template <typename A, int K, int SHAPE> struct TapViewV : TapView<A, K, SHAPE> {
    static constexpr bool kRowPipe = true;
    static constexpr bool kVgpr = true;
};

// the taps held in VGPRs for a whole tile (kTapResident); entries of skipped
// taps are never read and never materialised
template <typename A, int K, int SHAPE> struct TapRegs {
    static constexpr bool kRowPipe = false;
    static constexpr bool kVgpr = false;
    A t[K * K];
    static constexpr bool kSkip(int n, int c) { return TapView<A, K, SHAPE>::kSkip(n, c); }
    __device__ __forceinline__ A at(int n, int c) const { return t[n * K + c]; }
};

constexpr int kWavesPerBlock2D = 4;

// blockIdx -> tile-block so that XCD x (which receives blocks x, x+8, ...) works
// on one contiguous range of tile-blocks.
__device__ __forceinline__ int xcd_remap(int b, int nblocks) {
    const int n8 = (nblocks / 8) * 8;
    if (b >= n8) return b;
    return (b % 8) * (n8 / 8) + b / 8;
}

constexpr int cgcd(int a, int b) { return b == 0 ? a : cgcd(b, a % b); }

template <int R, int V, int T> struct Geom2D {
    static constexpr int K = 2 * R + 1;
    static constexpr int HL = (T * R + V - 1) / V;  // halo lanes per side
    static constexpr int OUT_LANES = kWave - 2 * HL;
    static constexpr int OW = OUT_LANES * V;        // output columns per strip
    static_assert(OUT_LANES > 0, "radius*steps too large for one wave");
};

// ------------------------------------------------------------------ one output row
// X[n][c]: rows y-r..y+r (n = 0..2r) of this lane's V columns.
// out[j] = sum_n sum_m w[n][m+r] * in(col j+m, row n) across lanes.
//
// Every tap (n, m) is read once (scalar load, uniform) and feeds its V FMAs at
// once, into the partial sum P[j][d] that belongs to lane distance
// d = floor((j+m)/V): d = 0 is this lane's own output, d > 0 is owed by the lane d
// to the right, d < 0 by the lane -d to the left.  Taps therefore stay live only
// briefly (large boxes do not run out of scalar registers).  The partials then
// travel systolically: a = P[DR]; a = shl(a) + P[DR-1]; ...; out += shl(a)
// (wave_shl:1, fused into v_add_f32_dpp), and the mirror image with wave_shr:1.
template <typename A, int R, int SHAPE, int V, typename TW>
__device__ __forceinline__ void stencil_row(const A (&X)[2 * R + 1][V], const TW& w, A (&out)[V]) {
    constexpr int K = 2 * R + 1;
    constexpr int D = (R + V - 1) / V;  // largest lane distance
    constexpr bool KAHAN = SHAPE == kBoxKahan;
    A P[V][2 * D + 1];
    A C[V][2 * D + 1];  // Kahan compensation (KAHAN only; dead code otherwise)
#pragma unroll
    for (int j = 0; j < V; ++j)
#pragma unroll
        for (int d = 0; d < 2 * D + 1; ++d) P[j][d] = C[j][d] = 0;
    // taps of the present filter row (row-pipelined forms; a star's rows other than
    // the centre row hold only the centre tap)
    auto row_taps = [&](int n, A (&t)[K]) {
#pragma unroll
        for (int c = 0; c < K; ++c) {
            if (SHAPE == kStar && n != R && c != R) continue;
            if (TW::kSkip(n, c)) continue;
            t[c] = w.at(n, c);
        }
    };
    A rt[K];
    if constexpr (TW::kRowPipe) row_taps(0, rt);
#pragma unroll
    for (int n = 0; n < K; ++n) {
        A nt[K];
        // VGPR copies of this row's taps come first (they wait for its loads), then
        // a barrier, so the next row's loads, issued below with this row's FMAs,
        // are not caught by that wait
        if constexpr (TW::kRowPipe && TW::kVgpr) {
#pragma unroll
            for (int c = 0; c < K; ++c) {
                if (SHAPE == kStar && n != R && c != R) continue;
                if (TW::kSkip(n, c)) continue;
                rt[c] = to_vgpr(rt[c]);
            }
        }
        bool issued = false;
        auto issue_next = [&]() {
            if constexpr (TW::kRowPipe) {
                if (!issued) {
                    issued = true;
                    __builtin_amdgcn_sched_barrier(0);
                    if (n + 1 < K) row_taps(n + 1, nt);
                }
            }
        };
        if constexpr (TW::kVgpr) issue_next();
        A Q[V][2 * D + 1];  // this filter row's partials (KAHAN, rows after the first)
        const bool fresh = KAHAN && n > 0;
#pragma unroll
        for (int j = 0; j < V; ++j)
#pragma unroll
            for (int d = 0; d < 2 * D + 1; ++d) Q[j][d] = 0;
#pragma unroll
        for (int m = -R; m <= R; ++m) {
            if (SHAPE == kStar && n != R && m != 0) continue;
            if (TW::kSkip(n, m + R)) continue;
            const A wt = TW::kRowPipe ? rt[m + R] : w.at(n, m + R);
#pragma unroll
            for (int j = 0; j < V; ++j) {
                const int c = j + m;
                const int d = c >= 0 ? c / V : -((-c + V - 1) / V);
                if (fresh)
                    Q[j][d + D] = fmaa(wt, X[n][c - d * V], Q[j][d + D]);
                else
                    P[j][d + D] = fmaa(wt, X[n][c - d * V], P[j][d + D]);
            }
            // SGPR taps: the first tap's FMAs wait for this row's loads; the next
            // row's loads go out after them
            issue_next();
        }
        if (fresh) {
#pragma unroll
            for (int j = 0; j < V; ++j)
#pragma unroll
                for (int d = 0; d < 2 * D + 1; ++d) {
                    // partials this row touched: lane distances floor((j-R)/V)..floor((j+R)/V)
                    const int lo = j - R >= 0 ? (j - R) / V : -((R - j + V - 1) / V);
                    const int hi = (j + R) / V;
                    if (d - D < lo || d - D > hi) continue;
                    const A y = Q[j][d] - C[j][d];
                    const A t = P[j][d] + y;
                    C[j][d] = (t - P[j][d]) - y;
                    P[j][d] = t;
                }
        }
        if constexpr (TW::kRowPipe) {
            // one filter row per scheduling region (a star's one-tap rows need none)
            if (SHAPE != kStar || n == R || n == R - 1) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int c = 0; c < K; ++c) rt[c] = nt[c];
        }
    }
    if (KAHAN) {
#pragma unroll
        for (int j = 0; j < V; ++j)
#pragma unroll
            for (int d = 0; d < 2 * D + 1; ++d) P[j][d] -= C[j][d];
    }
#pragma unroll
    for (int j = 0; j < V; ++j) {
        A acc = P[j][D];
        const int DR = (j + R) / V;               // farthest right lane with a partial
        const int DL = R > j ? (R - j + V - 1) / V : 0;
        if (DR >= 1) {
            A a = P[j][D + DR];
#pragma unroll
            for (int d = D; d >= 1; --d)
                if (d < DR) a = from_right(a) + P[j][D + d];
            acc += from_right(a);
        }
        if (DL >= 1) {
            A b = P[j][D - DL];
#pragma unroll
            for (int e = D; e >= 1; --e)
                if (e < DL) b = from_left(b) + P[j][D - e];
            acc += from_left(b);
        }
        out[j] = acc;
    }
}

// ------------------------------------------------------------------ kernel
// S storage type, R radius, SHAPE star/box, V columns per lane, T fused steps,
// PF prefetch distance in rows, RING ring-slot form, U ticks per unrolled block.
//
// Register rings.  Input row y lives in ring slot (y - (c0-2R)) mod NR0 where c0
// is the first consumed row; the rows of an intermediate step t produced at tick
// k live in slot (k + 2R) mod NRS.  Tick k loads row c0+k+PF and reads rows
// c0+k-2R .. c0+k (slots k .. k+2R).

//  * RING: the loop body is one full ring period (U = lcm(NR0, NRS) ticks), so
//    every slot index is a compile-time register and nothing ever moves; loads
//    stay in flight across ticks and blocks.
//  * !RING (large radius): U ticks per block; the block's U row loads are issued
//    at its start, and after the block the rings shift down by U slots with
//    register moves, which only touch rows whose loads had the whole block to
//    land.  Used where a full ring period would not fit the instruction cache.
// SO: the storage type of dst (default S).  A chain of fused launches over
// BF16 / FP16 grids keeps its running state in FP32 between launches: its first
// launch reads S and writes FP32, its last reads FP32 and writes S.
template <typename S, int R, int SHAPE, int V, int T, int PF, bool RING, int U, typename SO = S>
struct Stencil2D {
    using A = typename Storage<S>::acc;
    using G = Geom2D<R, V, T>;
    static constexpr int K = 2 * R + 1;
    static constexpr int NT = K * K;
    static constexpr int NR0 = RING ? K + PF : K + PF + U - 1;  // input ring
    static constexpr int NRS = RING ? K : K + U - 1;            // ring of each step
    static constexpr int NSTAGE_RINGS = T > 1 ? T - 1 : 1;
    static_assert(!RING || (U % NR0 == 0 && (T == 1 || U % NRS == 0)), "ring period");

    using Row = RowReg<S, V>;  // an input ring slot
    static constexpr int s0(int k, int n) { return RING ? (k + n) % NR0 : k + n; }
    static constexpr int sS(int k, int n) { return RING ? (k + n) % NRS : k + n; }

    struct Lane {
        const S* __restrict__ src;
        SO* __restrict__ dst;
        Params2D p;
        int x0, ys, ye;
        int ylast;      // last input row this tile needs (ye - 1 + T*R)
        uint32_t xoff;  // byte offset of this lane's strip within a src row (x0 >= 0)
        uint32_t xoff_out;  // ... within a dst row
        bool lane_full, edge, store_lane;
        bool inb[V];
    };

    // CHK = 0: the wave's whole footprint (rows and columns, halo included) is
    // inside the grid and vector-aligned, so rows load with one unconditional
    // vector access and no boundary rule applies.  CHK = 1: an edge tile under
    // zero padding (or HOLD); CHK = 2: an edge tile clamping to the edge.
    // Row base pointers are wave-uniform (SGPRs); a lane adds its 32-bit byte
    // offset, so loads and stores use the SGPR-base + VGPR-offset address form.
    __device__ static __forceinline__ const char* row_ptr(const Lane& L, int y) {
        return reinterpret_cast<const char*>(L.src) +
               static_cast<long long>(y) * L.p.pitch * static_cast<long long>(sizeof(S));
    }
    template <int CHK>
    __device__ static __forceinline__ void load_row(const Lane& L, int y, Row& dst) {
        if constexpr (!CHK) {
            dst.load_vec(reinterpret_cast<const S*>(row_ptr(L, y) + L.xoff));
        } else if constexpr (CHK == 1) {
            // Zero padding (and HOLD, whose stored cells never read outside the
            // grid): one raw-buffer access over the row, with no branch.  Rows
            // outside the grid take an offset past the buffer, columns outside
            // [0, width) lie past its end (x0 < 0 wraps past 2^31), and both read
            // 0 -- so the loads of an edge tile stay in flight across ticks like an
            // interior tile's (element loads under per-lane branches made the
            // compiler wait for every load at each row).
            const bool yin = y >= 0 && y < L.p.height;
            const BufRsrc rs = buf_rsrc(row_ptr(L, yin ? y : 0),
                                        static_cast<uint32_t>(L.p.width) * static_cast<uint32_t>(sizeof(S)));
            dst.load_buf(rs, yin ? L.xoff : kBufOOB);
            if constexpr (Row::kPacked) {
                // odd width: the dword holding the last element also holds the
                // element past the row end, so the range check drops it whole;
                // the last element comes on its own (0 for every other lane)
                const int k = L.p.width - 1 - L.x0;  // its index in this lane's strip
                const bool mine = yin && (L.p.width & 1) && k >= 0 && k < V;
                const uint32_t last = __builtin_amdgcn_raw_buffer_load_b16(
                    rs, static_cast<int>(mine ? L.xoff + static_cast<uint32_t>(k) * 2u : kBufOOB), 0, 0);
#pragma unroll
                for (int n = 0; n < Row::N; ++n) dst.r[n] = mine && k == 2 * n ? last : dst.r[n];
            }
        } else {
            // clamp to the edge: the nearest row, and per column the nearest
            // column for lanes that reach past an edge
            const char* row = row_ptr(L, clampi(y, 0, L.p.height - 1));
            if (L.lane_full && L.p.vec_ok) {
                dst.load_vec(reinterpret_cast<const S*>(row + L.xoff));
            } else {
                S e[V];
#pragma unroll
                for (int j = 0; j < V; ++j) {
                    const uint32_t o = static_cast<uint32_t>(clampi(L.x0 + j, 0, L.p.width - 1)) *
                                       static_cast<uint32_t>(sizeof(S));
                    e[j] = gld<S>(row + o);
                }
                dst.set(e);
            }
        }
    }

    // Interior tiles store without a branch: one raw buffer store over the output
    // row, whose offset is out of range (the hardware drops the store) for halo
    // lanes and for ticks whose row is outside the segment.  Branches around the
    // stores split the unrolled tick into separate scheduling regions (no
    // interleaving of the steps of a tick, no DPP-into-add fusion across them).
    __device__ static __forceinline__ void store_row_nb(const Lane& L, int y, const A (&v)[V]) {
        const bool ok = L.store_lane && y >= L.ys && y < L.ye;
        const int yy = y < 0 ? 0 : y;
        char* row = reinterpret_cast<char*>(L.dst) +
                    static_cast<long long>(yy) * L.p.pitch_out * static_cast<long long>(sizeof(SO));
        SO tmp[V];
#pragma unroll
        for (int j = 0; j < V; ++j) tmp[j] = Storage<SO>::store(v[j]);
        constexpr int W = V * static_cast<int>(sizeof(SO)) / 4;
        uint32_t words[W];
        __builtin_memcpy(words, tmp, sizeof(tmp));
        buf_store<uint32_t, W>(buf_rsrc(row, static_cast<uint32_t>(L.p.width) * sizeof(SO)),
                               ok ? L.xoff_out : kBufOOB, words);
    }

    // Edge tiles store through a raw buffer over the row with no branch: halo
    // lanes, rows outside the segment and lanes left of the grid take an offset
    // past the buffer; dwords past the row end are dropped by the range check,
    // which works per dword -- so a lane straddling the right edge stores exactly
    // its in-grid elements, except that a 2-byte element whose dword also holds
    // the element past the end (odd width) is stored on its own.  Keeping every
    // store unconditional keeps the memory-counter waits of the loads exact.
    __device__ static __forceinline__ void store_row_edge(const Lane& L, int y, const A (&v)[V]) {
        constexpr uint32_t ES = sizeof(SO);
        const bool ok = L.store_lane && y >= L.ys && y < L.ye;
        char* row = reinterpret_cast<char*>(L.dst) +
                    static_cast<long long>(y < 0 ? 0 : y) * L.p.pitch_out * static_cast<long long>(ES);
        const BufRsrc rs = buf_rsrc(row, static_cast<uint32_t>(L.p.width) * ES);
        SO tmp[V];
#pragma unroll
        for (int j = 0; j < V; ++j) tmp[j] = Storage<SO>::store(v[j]);
        constexpr int W = V * static_cast<int>(ES) / 4;
        uint32_t words[W];
        __builtin_memcpy(words, tmp, sizeof(tmp));
        buf_store<uint32_t, W>(rs, ok ? L.xoff_out : kBufOOB, words);
        if constexpr (ES == 2) {
            const int k = L.p.width - 1 - L.x0;  // the last column's index in this strip
            const bool mine = ok && (L.p.width & 1) && k >= 0 && k < V;
            uint16_t last = tmp[0].bits;
#pragma unroll
            for (int j = 1; j < V; ++j) last = k == j ? tmp[j].bits : last;
            __builtin_amdgcn_raw_buffer_store_b16(
                last, rs, static_cast<int>(mine ? L.xoff_out + static_cast<uint32_t>(k) * ES : kBufOOB), 0, 0);
        }
    }

    // interior tiles (scalar-tap kernels): one vector store per lane
    __device__ static __forceinline__ void store_row(const Lane& L, int y, const A (&v)[V]) {
        if (!L.store_lane) return;
        char* row = reinterpret_cast<char*>(L.dst) +
                    static_cast<long long>(y) * L.p.pitch_out * static_cast<long long>(sizeof(SO));
        store_strip_vec<SO, V>(reinterpret_cast<SO*>(row + L.xoff_out), v);
    }

    // BF16 / F16: cv holds the window rows converted to FP32, each row converted
    // once, at the tick where it enters the window (the packed ring slot keeps the
    // load's result until then); unused for F32 / F64.
    using TP = TapPolicy<A, R, SHAPE>;
    using TRegs = TapRegs<A, K, SHAPE>;

    // Pipeline fill.  Tick k of a tile (k = 0 consumes input row c0) runs step t
    // on a row the last step needs only once k >= 2(t-1)R: before that, step t's
    // rows lie above what step t+1 will ever read.  The first NPB blocks of a
    // tile are compiled with those (step, tick) pairs left out -- a triangle of
    // T(T-1)R step-ticks, 6-10% of a fused tile's work at T = 8 -- and compute
    // the same values as the full blocks for every row that is read (bit-equal
    // results).  Measured on 8192^2 FP32, old/new interleaved: T = 8 radius 1
    // 2-4% faster, radius 2 (V = 2) 3-7%; T = 4 radius 2 (V = 4) between 5%
    // slower and 4% faster, so fewer than 8 steps keep the plain loop.
    static constexpr int NPB = T >= 8 ? (2 * (T - 1) * R + U - 1) / U : 0;
    static constexpr bool skip_step(int pb, int kt, int t) { return pb >= 0 && pb * U + kt < 2 * (t - 1) * R; }

    // PB: the prologue block this tick belongs to (0 .. NPB-1), or -1 past them
    template <int CHK, int KT, int PB>
    __device__ static __forceinline__ void tick(const Lane& L, const A* __restrict__ w, const TRegs& tr,
                                                int yc0, Row (&ring0)[NR0], A (&cv)[NR0][V],
                                                A (&rings)[NSTAGE_RINGS][NRS][V]) {
        const int yc = yc0 + KT;
        // rows past the tile's last needed row only feed discarded ticks: re-read
        // that row instead (an L2 hit) so no HBM bytes are fetched for them
        if constexpr (RING) load_row<CHK>(L, min(yc + PF, L.ylast), ring0[s0(KT, PF + 2 * R)]);
        // non-resident taps come through scalar loads every tick; the opaque zero
        // offset keeps them from being hoisted out of the row loop, which large
        // boxes cannot afford
        int zero = 0;
        asm volatile("" : "+s"(zero));
        using CP = typename TapView<A, K, SHAPE>::CP;
        const TapView<A, K, SHAPE> tw{(CP)(w) + zero};
        const TapViewV<A, K, SHAPE> twv{{(CP)(w) + zero}};
        if constexpr (Row::kPacked) {
#pragma unroll
            for (int j = 0; j < V; ++j) cv[s0(KT, 2 * R)][j] = ring0[s0(KT, 2 * R)].get(j);
        }
#pragma unroll
        for (int t = 1; t <= T; ++t) {
            if (skip_step(PB, KT, t)) continue;  // a constant once the loop unrolls
            A X[K][V];
#pragma unroll
            for (int n = 0; n < K; ++n)
#pragma unroll
                for (int j = 0; j < V; ++j)
                    X[n][j] = (t == 1) ? (Row::kPacked ? cv[s0(KT, n)][j] : ring0[s0(KT, n)].get(j))
                                       : rings[t - 2][sS(KT, n)][j];
            A out[V];
            if constexpr (TP::mode == kTapResident) stencil_row<A, R, SHAPE, V>(X, tr, out);
            else if constexpr (TP::mode == kTapTick) stencil_row<A, R, SHAPE, V>(X, twv, out);
            else stencil_row<A, R, SHAPE, V>(X, tw, out);
            const int q = yc - t * R;  // grid row produced by this step
            if (t < T) {
                if constexpr (CHK) {
                    // zero outside the grid so the next step sees zero padding
                    const bool qin = q >= 0 && q < L.p.height;
#pragma unroll
                    for (int j = 0; j < V; ++j) out[j] = qin && L.inb[j] ? out[j] : A(0);
                }
#pragma unroll
                for (int j = 0; j < V; ++j) rings[t - 1][sS(KT, 2 * R)][j] = out[j];
            } else if constexpr (!CHK && TP::mode != kTapSgpr) {
                store_row_nb(L, q, out);
            } else if constexpr (CHK) {
                if constexpr (T == 1) {
                    // HOLD: border cells keep their input value
                    const bool hold = L.p.boundary == kHold;
                    const Row& in = ring0[s0(KT, R)];
                    const bool row_border = q < R || q >= L.p.height - R;
#pragma unroll
                    for (int j = 0; j < V; ++j) {
                        const int x = L.x0 + j;
                        const bool border = hold && (row_border || x < R || x >= L.p.width - R);
                        out[j] = border ? in.get(j) : out[j];
                    }
                }
                store_row_edge(L, q, out);
            } else if (q >= L.ys && q < L.ye) {
                store_row(L, q, out);
            }
        }
    }

    template <int CHK, int KT, int PB>
    __device__ static __forceinline__ void block(const Lane& L, const A* __restrict__ w, const TRegs& tr,
                                                 int yc0, Row (&ring0)[NR0], A (&cv)[NR0][V],
                                                 A (&rings)[NSTAGE_RINGS][NRS][V]) {
        if constexpr (KT < U) {
            tick<CHK, KT, PB>(L, w, tr, yc0, ring0, cv, rings);
            // Boxes with VGPR taps: one tick per scheduling region -- interleaving
            // whole ticks (free once the interior stores stopped branching) ran
            // past the register file.  Stars (few taps) interleave freely: a
            // barrier here cost the BF16 stars 20%.  SGPR taps keep the scheduler's
            // order, which hoists the next tick's tap loads (with a barrier here
            // they spilled).
            if constexpr (SHAPE != kStar && TP::mode != kTapSgpr) __builtin_amdgcn_sched_barrier(0);
            block<CHK, KT + 1, PB>(L, w, tr, yc0, ring0, cv, rings);
        }
    }

    // one block of U ticks: its row loads (!RING), the ticks, the ring shift (!RING)
    template <int CHK, int PB>
    __device__ static __forceinline__ void body(const Lane& L, const A* __restrict__ w, const TRegs& tr,
                                                int yc0, Row (&ring0)[NR0], A (&cv)[NR0][V],
                                                A (&rings)[NSTAGE_RINGS][NRS][V]) {
        if constexpr (!RING) {
#pragma unroll
            for (int k = 0; k < U; ++k)
                load_row<CHK>(L, min(yc0 + k + PF, L.ylast), ring0[2 * R + PF + k]);
        }
        block<CHK, 0, PB>(L, w, tr, yc0, ring0, cv, rings);
        if constexpr (!RING) {
#pragma unroll
            for (int i = 0; i < 2 * R + PF; ++i) ring0[i] = ring0[i + U];
            if constexpr (Row::kPacked) {
#pragma unroll
                for (int i = 0; i < 2 * R; ++i)
#pragma unroll
                    for (int j = 0; j < V; ++j) cv[i][j] = cv[i + U][j];
            }
#pragma unroll
            for (int s = 0; s < NSTAGE_RINGS; ++s)
#pragma unroll
                for (int i = 0; i < 2 * R; ++i)
#pragma unroll
                    for (int j = 0; j < V; ++j) rings[s][i][j] = rings[s][i + U][j];
        }
    }

    // the first NPB blocks of a tile (fewer when the tile is shorter)
    template <int CHK, int PB>
    __device__ static __forceinline__ void prologue(const Lane& L, const A* __restrict__ w, const TRegs& tr,
                                                    int& yc0, int& b, int nblk, Row (&ring0)[NR0],
                                                    A (&cv)[NR0][V], A (&rings)[NSTAGE_RINGS][NRS][V]) {
        if constexpr (PB < NPB) {
            if (b < nblk) {
                body<CHK, PB>(L, w, tr, yc0, ring0, cv, rings);
                ++b;
                yc0 += U;
                prologue<CHK, PB + 1>(L, w, tr, yc0, b, nblk, ring0, cv, rings);
            }
        }
    }

    template <int CHK>
    __device__ static __forceinline__ void stream(const Lane& L, const A* __restrict__ w, int c0,
                                                  int nblk) {
        Row ring0[NR0];
        A cv[NR0][V];
        A rings[NSTAGE_RINGS][NRS][V];
        TRegs tr;
        if constexpr (TP::mode == kTapResident) {
            const TapView<A, K, SHAPE> tv{(typename TapView<A, K, SHAPE>::CP)(w)};
#pragma unroll
            for (int n = 0; n < K; ++n)
#pragma unroll
                for (int c = 0; c < K; ++c) {
                    if (SHAPE == kStar && n != R && c != R) continue;
                    if (TRegs::kSkip(n, c)) continue;
                    tr.t[n * K + c] = to_vgpr(tv.at(n, c));
                }
        }
#pragma unroll
        for (int s = 0; s < NSTAGE_RINGS; ++s)
#pragma unroll
            for (int i = 0; i < NRS; ++i)
#pragma unroll
                for (int j = 0; j < V; ++j) rings[s][i][j] = 0;
#pragma unroll
        for (int i = 0; i < NR0; ++i) ring0[i].zero();
        // rows c0-2R .. c0+PF-1 into slots 0 .. 2R+PF-1
#pragma unroll
        for (int s = 0; s < 2 * R + PF; ++s) load_row<CHK>(L, min(c0 - 2 * R + s, L.ylast), ring0[s]);
        if constexpr (Row::kPacked) {
#pragma unroll
            for (int s = 0; s < 2 * R; ++s)
#pragma unroll
                for (int j = 0; j < V; ++j) cv[s][j] = ring0[s].get(j);
        }
        int yc0 = c0;
        int b = 0;
        // interior tiles only: with the edge paths' prologues too the kernel's
        // register count grew (R = 2, V = 4, T = 4: 144 -> 199, one wave less
        // per SIMD); edge tiles are few on the grids where fused steps pay
        if constexpr (CHK == 0) prologue<CHK, 0>(L, w, tr, yc0, b, nblk, ring0, cv, rings);
        for (; b < nblk; ++b, yc0 += U) body<CHK, -1>(L, w, tr, yc0, ring0, cv, rings);
    }

    // One tile: column strip `strip`, output rows [seg*seg_rows, ...) of the grid
    // described by p.  strip and seg are wave-uniform.
    __device__ static __forceinline__ void run_tile(const S* __restrict__ src, SO* __restrict__ dst,
                                                    const Params2D& p, const A* __restrict__ w,
                                                    int strip, int seg) {
        const int lane = threadIdx.x & (kWave - 1);
        Lane L;
        L.src = src;
        L.dst = dst;
        L.p = p;
        const int strip_x = strip * G::OW;
        L.x0 = strip_x + (lane - G::HL) * V;
        L.xoff = static_cast<uint32_t>(L.x0) * static_cast<uint32_t>(sizeof(S));
        L.xoff_out = static_cast<uint32_t>(L.x0) * static_cast<uint32_t>(sizeof(SO));
        L.ys = seg * p.seg_rows;
        L.ye = min(p.height, L.ys + p.seg_rows);
        L.lane_full = L.x0 >= 0 && L.x0 + V <= p.width;
        L.edge = (strip_x - G::HL * V < 0) || (strip_x - G::HL * V + kWave * V > p.width);
        L.store_lane = lane >= G::HL && lane < kWave - G::HL;
#pragma unroll
        for (int j = 0; j < V; ++j) L.inb[j] = (L.x0 + j >= 0) && (L.x0 + j < p.width);

        // first tick consumes input row c0; step T first emits row ys at tick ys+T*R;
        // the last tick emits row ye-1
        const int c0 = L.ys - T * R + 2 * R;
        const int ntick = (L.ye - 1 + T * R) - c0 + 1;
        const int nblk = (ntick + U - 1) / U;
        L.ylast = L.ye - 1 + T * R;
        const int y_lo = c0 - 2 * R;
        const int y_hi = min(c0 + nblk * U - 1 + PF, L.ylast);  // last row any tick loads
        // (interior stores address a row through a buffer of width*sizeof(SO) < 2^31
        // bytes, with 32-bit offsets)
        const bool interior = !L.edge && p.vec_ok && y_lo >= 0 && y_hi < p.height &&
                              static_cast<long long>(p.width) * static_cast<long long>(sizeof(SO)) < (1LL << 31);
        if (interior)
            stream<0>(L, w, c0, nblk);
        else if (p.boundary == kClamp)
            stream<2>(L, w, c0, nblk);
        else
            stream<1>(L, w, c0, nblk);
    }

    __device__ static void run(const S* __restrict__ src, SO* __restrict__ dst, const Params2D& p,
                               const A* __restrict__ w) {
        // wave index is uniform across the wave: keep the tile bookkeeping in SGPRs
        const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
        const int tile = xcd_remap(blockIdx.x, p.nblocks) * kWavesPerBlock2D + wv;
        if (tile >= p.ntiles) return;  // whole wave leaves; no barriers are used
        run_tile(src, dst, p, w, tile % p.nstrips, tile / p.nstrips);
    }
};

// The taps are the FIRST kernel argument, so they sit at offset 0 of the kernarg
// segment; the kernel reads them there with scalar loads instead of copying the
// by-value struct into private memory.
// Taps passed by value must leave room in the 4 KiB kernel-argument segment;
// larger sets (FP64 filters past 22 x 22) always come from device memory.
constexpr int kMaxKernargTapBytes = 3584;
template <typename A, int R, int SHAPE> struct KernargTaps {
    static constexpr int n = SHAPE == kBoxEven ? 4 * R * R : (2 * R + 1) * (2 * R + 1);
    static constexpr bool fits = n * static_cast<int>(sizeof(A)) <= kMaxKernargTapBytes;
    static constexpr int count = fits ? n : 1;
};

template <typename S, int R, int SHAPE, int V, int T, int PF, bool RING, int U, int WPE, typename SO = S>
__global__ __launch_bounds__(kWave * kWavesPerBlock2D) __attribute__((amdgpu_waves_per_eu(WPE)))
void ssam2d_kernel(
    Taps<typename Storage<S>::acc, KernargTaps<typename Storage<S>::acc, R, SHAPE>::count> tp,
    const S* __restrict__ src, SO* __restrict__ dst, Params2D p) {
    (void)tp;
    using A = typename Storage<S>::acc;
    static_assert(std::is_same<A, typename Storage<SO>::acc>::value, "one accumulation type");
    const A* __restrict__ w = p.taps ? static_cast<const A*>(p.taps)
                                     : (const A*)(__builtin_amdgcn_kernarg_segment_ptr());
    Stencil2D<S, R, SHAPE, V, T, PF, RING, U, SO>::run(src, dst, p, w);
}

}  // namespace ssam
