// stencil2d_lds.hpp -- the LDS-tap form of the 2D box / dense-filter kernel
// (FP32 accumulation), for the large filters of
//   stencil2d_64pt/81/121pt::j2d*pt (ssai-2dstencil/stencil2D_{64,81,121}pt.cu)
//   CONV2D::kernel_convolution2D     (ssai-2dconv/conv2D.cu:9-70)
//
// Why a second form.  gfx950 issues v_fmac_f32 at twice the single rate only when
// every operand is a VGPR and at least two waves per SIMD have FMAs ready (a
// scalar tap operand, or one wave, leaves the single rate).  The register-cache
// kernel (stencil2d.hpp) therefore keeps a box's taps resident in VGPRs, which
// for 81 / 121 taps leaves room for only two waves per SIMD, and larger filters
// fall back to scalar taps.  Here the workgroup stages the taps in LDS once; each
// wave reads one filter row at a time (16-byte broadcast reads, uniform address,
// no bank conflicts) into a double-buffered VGPR row, and every tap so read feeds
// M output rows x V columns of FMAs.  The input rows still stream through the
// register ring of the register-cache kernel (each row loaded once per tile), the
// partial sums still move lane to lane by DPP -- only the tap storage changes,
// which frees about K*K VGPRs for more waves per SIMD.
//
// Output rows come M per block: block b reads input rows yc0-2R .. yc0+M-1 (ring
// slots 0 .. 2R+M-1) and emits rows yc0-R .. yc0-R+M-1; the block's M new rows
// are loaded PF rows ahead at its start, and the ring shifts down by M after it.
#pragma once
#include "stencil2d.hpp"

namespace ssam {

template <typename A> struct LdsQuad;
template <> struct LdsQuad<float> {
    static constexpr int N = 4;
    typedef float type __attribute__((ext_vector_type(4)));
};
template <> struct LdsQuad<double> {
    static constexpr int N = 2;
    typedef double type __attribute__((ext_vector_type(2)));
};

// Taps in LDS: K rows of KC elements (K rounded up to whole 16-byte reads), row n
// holding w[n][0..K); for kBoxEven the last row and column are zero and unread.
template <typename A, int K> struct LdsTaps {
    static constexpr int N = LdsQuad<A>::N;
    static constexpr int NQ = (K + N - 1) / N;  // 16-byte reads per filter row
    static constexpr int KC = NQ * N;
    static constexpr int size = K * KC;
};

template <typename A> using LdsPtr = const __attribute__((address_space(3))) A*;

template <typename A, int K>
__device__ __forceinline__ void lds_tap_row(LdsPtr<A> tl, int n, A (&t)[LdsTaps<A, K>::KC]) {
    using Q = LdsQuad<A>;
    using LT = LdsTaps<A, K>;
    using QP = const __attribute__((address_space(3))) typename Q::type*;
    const QP p = reinterpret_cast<QP>(tl + n * LT::KC);
#pragma unroll
    for (int q = 0; q < LT::NQ; ++q) {
        const typename Q::type v = p[q];
#pragma unroll
        for (int e = 0; e < Q::N; ++e) t[q * Q::N + e] = v[e];
    }
}

// The systolic hop of one output row: P[j][d + D] holds the partial this lane
// owes the lane at distance d for column j (see stencil_row in stencil2d.hpp).
template <typename A, int R, int V>
__device__ __forceinline__ void hop_row(const A (&P)[V][2 * ((R + V - 1) / V) + 1], A (&out)[V]) {
    constexpr int D = (R + V - 1) / V;
#pragma unroll
    for (int j = 0; j < V; ++j) {
        A acc = P[j][D];
        const int DR = (j + R) / V;
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

// M output rows at once.  X[s]: input rows y0-R+s (s = 0 .. 2R+M-1) of this lane's
// V columns; out[i] = row y0+i.  Tap row n is read from LDS while row n-1
// computes, and used for all M rows.
template <typename A, int R, int SHAPE, int V, int M>
__device__ __forceinline__ void stencil_rows_lds(const A (&X)[2 * R + M][V], LdsPtr<A> tl, A (&out)[M][V]) {
    constexpr int K = 2 * R + 1;
    constexpr int D = (R + V - 1) / V;
    using LT = LdsTaps<A, K>;
    using TV = TapView<A, K, SHAPE>;
    // a kBoxEven filter's last row holds no taps
    constexpr int KR = SHAPE == kBoxEven ? K - 1 : K;
    A P[M][V][2 * D + 1];
#pragma unroll
    for (int i = 0; i < M; ++i)
#pragma unroll
        for (int j = 0; j < V; ++j)
#pragma unroll
            for (int d = 0; d < 2 * D + 1; ++d) P[i][j][d] = 0;
    A tb[2][LT::KC];
    lds_tap_row<A, K>(tl, 0, tb[0]);
    // Each filter row ends with an empty volatile asm on every partial: the row's
    // FMAs cannot sink past it and the next row's cannot rise above it, and the
    // tap reads (ordered with the asm statements) go out one row ahead.  Without
    // it instruction selection sinks all FMAs below all tap reads, so every tap
    // row is live at once and the registers this form frees are lost again.
#pragma unroll
    for (int n = 0; n < KR; ++n) {
        if (n + 1 < KR) lds_tap_row<A, K>(tl, n + 1, tb[(n + 1) & 1]);
#pragma unroll
        for (int m = -R; m <= R; ++m) {
            if (TV::kSkip(n, m + R)) continue;
            const A wt = tb[n & 1][m + R];
#pragma unroll
            for (int i = 0; i < M; ++i)
#pragma unroll
                for (int j = 0; j < V; ++j) {
                    const int c = j + m;
                    const int d = c >= 0 ? c / V : -((-c + V - 1) / V);
                    P[i][j][d + D] = fmaa(wt, X[n + i][c - d * V], P[i][j][d + D]);
                }
        }
#pragma unroll
        for (int i = 0; i < M; ++i)
#pragma unroll
            for (int j = 0; j < V; ++j) {
                // partials of column j: lane distances floor((j-R)/V) .. floor((j+R)/V)
                const int lo = j - R >= 0 ? (j - R) / V : -((R - j + V - 1) / V);
                const int hi = (j + R) / V;
#pragma unroll
                for (int d = lo; d <= hi; ++d) asm volatile("" : "+v"(P[i][j][d + D]));
            }
    }
#pragma unroll
    for (int i = 0; i < M; ++i) hop_row<A, R, V>(P[i], out[i]);
}

// S storage (F32 / F64), R radius, SHAPE kBox or kBoxEven, V columns per lane,
// M output rows per block, PF prefetch distance in rows (PF >= M).
template <typename S, int R, int SHAPE, int V, int M, int PF>
struct Stencil2DLds : Stencil2D<S, R, SHAPE, V, 1, PF, false, M> {
    using Base = Stencil2D<S, R, SHAPE, V, 1, PF, false, M>;
    using A = typename Storage<S>::acc;
    using G = Geom2D<R, V, 1>;
    using Row = typename Base::Row;
    using Lane = typename Base::Lane;
    static constexpr int K = 2 * R + 1;
    static constexpr int NR0 = K + PF + M - 1;  // input ring
    static_assert(PF >= M, "a block's new rows are consumed by later blocks");

    template <int CHK>
    __device__ static __forceinline__ void block(const Lane& L, LdsPtr<A> tl, int yc0, Row (&ring0)[NR0]) {
        A X[2 * R + M][V];
#pragma unroll
        for (int s = 0; s < 2 * R + M; ++s)
#pragma unroll
            for (int j = 0; j < V; ++j) X[s][j] = ring0[s].get(j);
        // the tap reads are the same in every block: an opaque zero offset keeps
        // the compiler from hoisting all K*K of them out of the row loop (they
        // would stay live in VGPRs, the register pressure this form removes)
        int zero = 0;
        asm volatile("" : "+v"(zero));
        A out[M][V];
        stencil_rows_lds<A, R, SHAPE, V, M>(X, tl + zero, out);
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const int q = yc0 + i - R;
            if constexpr (!CHK) {
                Base::store_row_nb(L, q, out[i]);
            } else {
                // HOLD: border cells keep their input value (ring slot of row q)
                const bool hold = L.p.boundary == kHold;
                const bool row_border = q < R || q >= L.p.height - R;
#pragma unroll
                for (int j = 0; j < V; ++j) {
                    const int x = L.x0 + j;
                    const bool border = hold && (row_border || x < R || x >= L.p.width - R);
                    out[i][j] = border ? X[R + i][j] : out[i][j];
                }
                Base::store_row_edge(L, q, out[i]);
            }
        }
    }

    template <int CHK>
    __device__ static __forceinline__ void stream(const Lane& L, LdsPtr<A> tl, int c0, int nblk) {
        Row ring0[NR0];
#pragma unroll
        for (int i = 0; i < NR0; ++i) ring0[i].zero();
#pragma unroll
        for (int s = 0; s < 2 * R + PF; ++s)
            Base::template load_row<CHK>(L, min(c0 - 2 * R + s, L.ylast), ring0[s]);
        int yc0 = c0;
        for (int b = 0; b < nblk; ++b, yc0 += M) {
#pragma unroll
            for (int k = 0; k < M; ++k)
                Base::template load_row<CHK>(L, min(yc0 + k + PF, L.ylast), ring0[2 * R + PF + k]);
            block<CHK>(L, tl, yc0, ring0);
#pragma unroll
            for (int i = 0; i < 2 * R + PF; ++i) ring0[i] = ring0[i + M];
        }
    }

    __device__ static __forceinline__ void run_tile(const S* __restrict__ src, S* __restrict__ dst,
                                                    const Params2D& p, LdsPtr<A> tl, int strip, int seg) {
        const int lane = threadIdx.x & (kWave - 1);
        Lane L;
        L.src = src;
        L.dst = dst;
        L.p = p;
        const int strip_x = strip * G::OW;
        L.x0 = strip_x + (lane - G::HL) * V;
        L.xoff = static_cast<uint32_t>(L.x0) * static_cast<uint32_t>(sizeof(S));
        L.xoff_out = L.xoff;
        L.ys = seg * p.seg_rows;
        L.ye = min(p.height, L.ys + p.seg_rows);
        L.lane_full = L.x0 >= 0 && L.x0 + V <= p.width;
        L.edge = (strip_x - G::HL * V < 0) || (strip_x - G::HL * V + kWave * V > p.width);
        L.store_lane = lane >= G::HL && lane < kWave - G::HL;
#pragma unroll
        for (int j = 0; j < V; ++j) L.inb[j] = (L.x0 + j >= 0) && (L.x0 + j < p.width);
        // block b consumes rows up to c0 + b*M + M - 1 and emits rows c0-R+b*M ...
        const int c0 = L.ys + R;
        const int ntick = L.ye - L.ys;
        const int nblk = (ntick + M - 1) / M;
        L.ylast = L.ye - 1 + R;
        const int y_lo = c0 - 2 * R;
        const int y_hi = min(c0 + nblk * M - 1 + PF, L.ylast);
        const bool interior = !L.edge && p.vec_ok && y_lo >= 0 && y_hi < p.height &&
                              static_cast<long long>(p.width) * static_cast<long long>(sizeof(S)) < (1LL << 31);
        if (interior)
            stream<0>(L, tl, c0, nblk);
        else if (p.boundary == kClamp)
            stream<2>(L, tl, c0, nblk);
        else
            stream<1>(L, tl, c0, nblk);
    }
};

template <typename S, int R, int SHAPE, int V, int M, int PF, int WPE>
__global__ __launch_bounds__(kWave * kWavesPerBlock2D) __attribute__((amdgpu_waves_per_eu(WPE)))
void ssam2d_lds_kernel(
    Taps<typename Storage<S>::acc, KernargTaps<typename Storage<S>::acc, R, SHAPE>::count> tp,
    const S* __restrict__ src, S* __restrict__ dst, Params2D p) {
    (void)tp;
    using A = typename Storage<S>::acc;
    constexpr int K = 2 * R + 1;
    using LT = LdsTaps<A, K>;
    using TV = TapView<A, K, SHAPE>;
    __shared__ __attribute__((aligned(16))) A tl[LT::size];
    const A* __restrict__ w = p.taps ? static_cast<const A*>(p.taps)
                                     : (const A*)(__builtin_amdgcn_kernarg_segment_ptr());
    for (int i = threadIdx.x; i < LT::size; i += kWave * kWavesPerBlock2D) {
        const int n = i / LT::KC, c = i - n * LT::KC;
        A v = 0;
        if (c < K && !TV::kSkip(n, c)) v = SHAPE == kBoxEven ? w[n * (K - 1) + c] : w[n * K + c];
        tl[i] = v;
    }
    __syncthreads();
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int tile = xcd_remap(blockIdx.x, p.nblocks) * kWavesPerBlock2D + wv;
    if (tile >= p.ntiles) return;
    Stencil2DLds<S, R, SHAPE, V, M, PF>::run_tile(src, dst, p, (LdsPtr<A>)(tl),
                                                  tile % p.nstrips, tile / p.nstrips);
}

}  // namespace ssam
