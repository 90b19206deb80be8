// stencil3d_hlt.hpp -- S sweeps of a radius-1 3D stencil per launch (S >= 1):
// the temporal blocking of Stencil3DHL2 carried to more sweeps, so a multi-sweep
// run reads and writes the grid once per S sweeps (the reference's IT fused
// steps, ssai-3dstencil/tmp3d7pt-v2.cu:62-208, generalised).
//
// A wave owns an x-strip (HL = ceil(S/V) halo lanes per side, or HLX if more,
// whose columns only feed their neighbours) and a y-block of NY output rows, and
// streams z.  (HLX lets passes of fewer sweeps share the tiles of a longer one.)
// Sweep k (k = 1..S) works one plane ahead of sweep k+1: at tick q it forms
// plane q+S-k over rows y0-(S-k) .. y0+NY-1+(S-k) from the three neighbouring
// planes of sweep k-1 (sweep 0 = the input ring of NY+2S rows), and keeps it in
// a three-plane ring; sweep S writes plane q.  Every intermediate value outside
// the grid's interior is 0, as a separate sweep writes 0 there, and each point
// sums its taps in the order of the one-sweep kernels, so S sweeps here equal S
// one-sweep launches bit for bit.
//
// Addressing as Stencil3DHL (one raw buffer, grid < 2 GiB): rows, columns and
// planes outside the grid read 0 or land on other rows of the grid, and either
// way only reach values that the interior rule sets to 0.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "ssam_device.hpp"
#include "stencil3d_hl.hpp"

namespace ssam {

template <typename SI, int SHAPE, int V, int NY, int PF, typename SO, int S, int HLX = 0>
struct Stencil3DHLT {
    using A = typename Storage<SI>::acc;
    static_assert(std::is_same<A, typename Storage<SO>::acc>::value, "one accumulation type");
    static_assert(PF >= 1 && S >= 1, "prefetch at least one plane; one sweep or more");
    static constexpr int NT = ntaps3(SHAPE, 1);
    static constexpr int HL = (S + V - 1) / V > HLX ? (S + V - 1) / V : HLX;  // halo lanes per side
    static constexpr int OW = (kWave - 2 * HL) * V;  // output columns per strip
    static constexpr int NI = NY + 2 * S;            // input rows y0-S .. y0+NY-1+S
    static constexpr int NM = NY + 2 * (S - 1);      // rows of sweep 1 (sweep k uses NY+2(S-k))
    static constexpr int NMR = S > 1 ? S - 1 : 1;    // intermediate rings (none used when S = 1)
    static constexpr int NPI = 3 + PF;               // input planes in the ring
    static constexpr int PERIOD = clcm(NPI, 3);      // ticks until all rings realign
    static constexpr uint32_t ESI = sizeof(SI), ESO = sizeof(SO);

    struct PlaneI { A v[NI][V]; };
    struct PlaneM { A v[NM][V]; };

    struct Ctx {
        BufRsrc src, dst;
        int ny, nz, zs;
        uint32_t syi, szi, syo, szo;  // byte strides of src and dst
        int y0;
        uint32_t lo;
        uint32_t so[V];
        bool xin[V];
    };

    template <bool VEC>
    __device__ static __forceinline__ void load_plane(const Ctx& c, int z, PlaneI& P) {
        const uint32_t pz = static_cast<uint32_t>(z) * c.szi + static_cast<uint32_t>(c.y0 - S) * c.syi;
#pragma unroll
        for (int i = 0; i < NI; ++i)
            buf_load_elems<SI, V, VEC>(c.src, c.lo + pz + static_cast<uint32_t>(i) * c.syi, P.v[i]);
    }

    template <int NR>
    __device__ static __forceinline__ A at(const A (&v)[NR][V], int i, int j, int dx) {
        const int cc = j + dx;
        if (cc >= 0 && cc < V) return v[i][cc];
        if (cc < 0) return from_left(v[i][V - 1]);
        return from_right(v[i][0]);
    }

    // tick KT of a period starting at q = zq0 (zq0 - (zs - 2(S-1)) is a multiple
    // of PERIOD): input plane zq0+k sits in slot k % NPI, and sweep k's plane
    // zq0+j in slot j % 3 of mid[k-1].
    template <bool VEC, int KT>
    __device__ static __forceinline__ void tick(const Ctx& c, const A* __restrict__ w, int zq0,
                                                PlaneI (&in)[NPI], PlaneM (&mid)[NMR][3]) {
        const int q = zq0 + KT;
        load_plane<VEC>(c, q + S + PF, in[(KT + S + PF) % NPI]);
        int zero = 0;
        asm volatile("" : "+s"(zero));
        const A* __restrict__ tw = w + zero;
#pragma unroll
        for (int k = 1; k < S; ++k) {
            const int m = q + S - k;  // plane of sweep k
            const bool zin = m >= 1 && m < c.nz - 1;
            PlaneM& M = mid[k - 1][(KT + S - k) % 3];
#pragma unroll
            for (int i = 0; i < NY + 2 * (S - k); ++i) {
                const int y = c.y0 - (S - k) + i;
                const bool yin = zin && y >= 1 && y < c.ny - 1;
#pragma unroll
                for (int j = 0; j < V; ++j) {
                    A acc = 0;
#pragma unroll
                    for (int t = 0; t < NT; ++t) {
                        const Tap3 tp = tap3(SHAPE, t);
                        const A x = k == 1 ? at(in[(KT + S - 1 + tp.dz) % NPI].v, i + 1 + tp.dy, j, tp.dx)
                                           : at(mid[k - 2][(KT + S - k + tp.dz + 3) % 3].v, i + 1 + tp.dy, j, tp.dx);
                        acc = fmaa(tw[t], x, acc);
                    }
                    M.v[i][j] = (yin && c.xin[j]) ? acc : A(0);
                }
            }
        }
        // sweep S: output plane q, rows y0 .. y0+NY-1
        if (q >= c.zs) {
            const bool zin = q >= 1 && q < c.nz - 1;
            const uint32_t qz = static_cast<uint32_t>(q) * c.szo;
#pragma unroll
            for (int i = 0; i < NY; ++i) {
                const int y = c.y0 + i;
                const bool yin = zin && y >= 1 && y < c.ny - 1;
                A out[V];
#pragma unroll
                for (int j = 0; j < V; ++j) {
                    A acc = 0;
#pragma unroll
                    for (int t = 0; t < NT; ++t) {
                        const Tap3 tp = tap3(SHAPE, t);
                        const A x = S == 1 ? at(in[(KT + tp.dz + NPI) % NPI].v, i + 1 + tp.dy, j, tp.dx)
                                           : at(mid[S > 1 ? S - 2 : 0][(KT + tp.dz + 3) % 3].v, i + 1 + tp.dy, j, tp.dx);
                        acc = fmaa(tw[t], x, acc);
                    }
                    out[j] = (yin && c.xin[j]) ? acc : A(0);
                }
                if (y < c.ny) buf_store_elems<SO, V, VEC>(c.dst, c.so, qz + static_cast<uint32_t>(y) * c.syo, out);
            }
        }
    }

    template <bool VEC, int KT>
    __device__ static __forceinline__ void period(const Ctx& c, const A* __restrict__ w, int zq0,
                                                  PlaneI (&in)[NPI], PlaneM (&mid)[NMR][3]) {
        if constexpr (KT < PERIOD) {
            tick<VEC, KT>(c, w, zq0, in, mid);
            period<VEC, KT + 1>(c, w, zq0, in, mid);
        }
    }

    template <bool VEC, int KT>
    __device__ static __forceinline__ void tail(const Ctx& c, const A* __restrict__ w, int zq0, int zend,
                                                PlaneI (&in)[NPI], PlaneM (&mid)[NMR][3]) {
        if constexpr (KT < PERIOD - 1) {
            if (zq0 + KT < zend) {
                tick<VEC, KT>(c, w, zq0, in, mid);
                tail<VEC, KT + 1>(c, w, zq0, zend, in, mid);
            }
        }
    }

    template <bool VEC>
    __device__ static __forceinline__ void stream(const Ctx& c, const A* __restrict__ w, int zs, int ze) {
        PlaneI in[NPI];
        PlaneM mid[NMR][3];
        // the first ticks read intermediate planes no sweep has formed yet (their
        // results feed only planes that are never stored): start them at 0
#pragma unroll
        for (int k = 0; k < NMR; ++k)
#pragma unroll
            for (int s = 0; s < 3; ++s)
#pragma unroll
                for (int i = 0; i < NM; ++i)
#pragma unroll
                    for (int j = 0; j < V; ++j) mid[k][s].v[i][j] = A(0);
        // ticks run q = zs-2(S-1) .. ze-1; the first tick needs input planes
        // q+S-2 .. q+S+PF-1 (slots S-2 .. S+PF-1, mod NPI)
        const int zq = zs - 2 * (S - 1);
#pragma unroll
        for (int k = S - 2; k < S + PF; ++k) load_plane<VEC>(c, zq + k, in[(k + NPI) % NPI]);
        int zq0 = zq;
#pragma nounroll
        for (; zq0 + PERIOD <= ze; zq0 += PERIOD) period<VEC, 0>(c, w, zq0, in, mid);
        tail<VEC, 0>(c, w, zq0, ze, in, mid);
    }

    // One tile: strip `strip`, y-block `yb`, output planes [zs, ze) of a dense
    // grid that fits grid3d_fits_buffer, S sweeps from src to dst.
    __device__ static __forceinline__ void run_tile(const SI* __restrict__ src, SO* __restrict__ dst,
                                                    const Grid3D& g, bool vec, const A* __restrict__ w,
                                                    int strip, int yb, int zs, int ze) {
        const int lane = threadIdx.x & (kWave - 1);
        const uint32_t n = static_cast<uint32_t>(g.nx) * static_cast<uint32_t>(g.ny) * static_cast<uint32_t>(g.nz);
        Ctx c;
        c.src = buf_rsrc(src, n * ESI);
        c.dst = buf_rsrc(dst, n * ESO);
        c.ny = g.ny;
        c.nz = g.nz;
        c.zs = zs;
        c.syi = static_cast<uint32_t>(g.sy) * ESI;
        c.szi = static_cast<uint32_t>(g.sz) * ESI;
        c.syo = static_cast<uint32_t>(g.sy) * ESO;
        c.szo = static_cast<uint32_t>(g.sz) * ESO;
        c.y0 = yb * NY;
        const int x0 = strip * OW + (lane - HL) * V;
        c.lo = static_cast<uint32_t>(x0) * ESI;
        const bool store_lane = lane >= HL && lane < kWave - HL;
#pragma unroll
        for (int j = 0; j < V; ++j) {
            const int x = x0 + j;
            c.so[j] = (store_lane && x < g.nx) ? static_cast<uint32_t>(x) * ESO : kBufOOB;
            c.xin[j] = x >= 1 && x < g.nx - 1;
        }
        if (V == 1 || vec) stream<true>(c, w, zs, ze);
        else stream<false>(c, w, zs, ze);
    }
};

}  // namespace ssam
