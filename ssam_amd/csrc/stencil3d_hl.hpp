// stencil3d_hl.hpp -- radius-1 3D stencils for narrow grids (tens to a few
// hundred columns), the tile body of the batched 3D launch.
//
// Same register-cache scheme as stencil3d.hpp (x-strip per wave, NY rows per
// lane, z streamed through a ring of planes, x-neighbours over DPP) with one
// change of layout that suits narrow grids: the strip's outer lanes are halo
// lanes.  Lane l of strip s holds columns x0 = s*OW + (l-1)*V .. x0+V-1 with
// OW = 62*V; lanes 0 and 63 only feed their neighbours and store nothing.
//
// Addressing: the grid is one raw buffer (< 2 GiB) and every access is a 32-bit
// byte offset, the row part wave-uniform, the column part per lane.  A row or
// column outside the grid gives an offset that either wraps past the buffer end
// (the load reads 0) or lands on another row of the grid; either way the value
// only reaches outputs of the boundary layer, which are written as 0.  So the
// hot loop has no clamp, no edge branch and no 64-bit address arithmetic; a lane
// that stores nothing gets an out-of-range store offset, and only the rows of the
// last y-block below the grid skip their store (a wave-uniform branch).
//
// Output rule (zero Dirichlet, as stencil3d.hpp): interior points get the
// stencil, the outer layer of the grid gets 0.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "ssam_device.hpp"
#include "stencil3d.hpp"

namespace ssam {

constexpr int cgcd3(int a, int b) { return b == 0 ? a : cgcd3(b, a % b); }

struct Grid3D {
    int nx, ny, nz;
    long long sy, sz;  // element strides of y and z (dense: nx, nx*ny)
};

// Whether a grid fits the 32-bit byte offsets of the batched tile.
inline bool grid3d_fits_buffer(long long nx, long long ny, long long nz, int elem_bytes) {
    return nx * ny * nz * elem_bytes < (1LL << 31) - 4096;
}

// SI: storage of src; SO: storage of dst (default SI).  A chain of launches over
// BF16 / FP16 grids keeps its running state in FP32: S -> FP32, FP32 -> FP32 and
// FP32 -> S passes.  Accumulation in Storage<SI>::acc (FP32, or FP64 for FP64).
template <typename SI, int SHAPE, int V, int NY, int PF, typename SO = SI>
struct Stencil3DHL {
    using A = typename Storage<SI>::acc;
    static_assert(std::is_same<A, typename Storage<SO>::acc>::value, "one accumulation type");
    static constexpr int R = 1;
    static constexpr int NT = ntaps3(SHAPE, R);
    static constexpr int NROW = NY + 2;     // rows cached per plane
    static constexpr int NP = 3 + PF;       // planes in the ring
    static constexpr int OW = (kWave - 2) * V;
    static constexpr uint32_t ESI = sizeof(SI), ESO = sizeof(SO);

    struct Plane { A v[NROW][V]; };

    struct Ctx {
        BufRsrc src, dst;
        int ny, nz;
        uint32_t syi, szi;    // byte strides of src
        uint32_t syo, szo;    // ... of dst
        int y0;
        uint32_t lo;          // byte offset of the lane's first column in its src row
        uint32_t so[V];       // dst offset of each column in its row (kBufOOB: not stored)
        bool xin[V];          // interior column
    };

    // VEC: rows are V-aligned (nx % V == 0), one V-wide access per row;
    // otherwise V single-element accesses.
    template <bool VEC>
    __device__ static __forceinline__ void load_plane(const Ctx& c, int z, Plane& P) {
        const uint32_t pz = static_cast<uint32_t>(z) * c.szi + static_cast<uint32_t>(c.y0 - 1) * c.syi;
#pragma unroll
        for (int i = 0; i < NROW; ++i)
            buf_load_elems<SI, V, VEC>(c.src, c.lo + pz + static_cast<uint32_t>(i) * c.syi, P.v[i]);
    }

    __device__ static __forceinline__ A at(const Plane& P, int i, int j, int dx) {
        const int cc = j + dx;
        if (cc >= 0 && cc < V) return P.v[i][cc];
        if (cc < 0) return from_left(P.v[i][V - 1]);
        return from_right(P.v[i][0]);
    }

    template <bool VEC, int KT>
    __device__ static __forceinline__ void tick(const Ctx& c, const A* __restrict__ w, int zq0,
                                                Plane (&ring)[NP]) {
        const int q = zq0 + KT;  // output plane
        load_plane<VEC>(c, q + 1 + PF, ring[(KT + 2 + PF) % NP]);
        int zero = 0;
        asm volatile("" : "+s"(zero));
        const A* __restrict__ tw = w + zero;
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
                    const Plane& P = ring[(KT + 1 + tp.dz + NP) % NP];
                    acc = fmaa(tw[t], at(P, i + 1 + tp.dy, j, tp.dx), acc);
                }
                out[j] = (yin && c.xin[j]) ? acc : A(0);
            }
            if (y < c.ny) buf_store_elems<SO, V, VEC>(c.dst, c.so, qz + static_cast<uint32_t>(y) * c.syo, out);
        }
    }

    // A full ring period (NP ticks) as straight-line code: the compiler's memory
    // counter waits stay exact, so prefetched planes stay in flight across it.
    template <bool VEC, int KT>
    __device__ static __forceinline__ void period(const Ctx& c, const A* __restrict__ w, int zq0,
                                                  Plane (&ring)[NP]) {
        if constexpr (KT < NP) {
            tick<VEC, KT>(c, w, zq0, ring);
            period<VEC, KT + 1>(c, w, zq0, ring);
        }
    }

    // the last partial period: ticks zq0+KT < zend only
    template <bool VEC, int KT>
    __device__ static __forceinline__ void tail(const Ctx& c, const A* __restrict__ w, int zq0, int zend,
                                                Plane (&ring)[NP]) {
        if constexpr (KT < NP - 1) {
            if (zq0 + KT < zend) {
                tick<VEC, KT>(c, w, zq0, ring);
                tail<VEC, KT + 1>(c, w, zq0, zend, ring);
            }
        }
    }

    template <bool VEC>
    __device__ static __forceinline__ void stream(const Ctx& c, const A* __restrict__ w, int zs, int ze) {
        Plane ring[NP];
        // planes zs-1 .. zs+PF into the slots of the first period
#pragma unroll
        for (int k = 0; k < 2 + PF; ++k) load_plane<VEC>(c, zs - 1 + k, ring[k]);
        int zq0 = zs;
#pragma nounroll
        for (; zq0 + NP <= ze; zq0 += NP) period<VEC, 0>(c, w, zq0, ring);
        tail<VEC, 0>(c, w, zq0, ze, ring);
    }

    // One tile: strip `strip`, y-block `yb`, output planes [zs, ze) (wave-uniform)
    // of a dense grid (sy = nx, sz = nx*ny) that fits grid3d_fits_buffer.
    // vec: nx % V == 0 and src/dst aligned to V elements.
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
        c.syi = static_cast<uint32_t>(g.sy) * ESI;
        c.szi = static_cast<uint32_t>(g.sz) * ESI;
        c.syo = static_cast<uint32_t>(g.sy) * ESO;
        c.szo = static_cast<uint32_t>(g.sz) * ESO;
        c.y0 = yb * NY;
        const int x0 = strip * OW + (lane - 1) * V;
        c.lo = static_cast<uint32_t>(x0) * ESI;  // x0 = -V wraps: reads past the end
        const bool store_lane = lane >= 1 && lane < kWave - 1;
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

constexpr int clcm(int a, int b) { return a / cgcd3(a, b) * b; }

// Two sweeps per pass (temporal blocking, the reference's IT = 2): the wave
// streams z once and keeps two register rings, input planes (NY+4 rows: the
// second sweep's halo needs two rows above and below) and the first sweep's
// planes (NY+2 rows, three planes).  Tick q loads input plane q+2+PF, computes
// first-sweep plane q+1 from input planes q..q+2 and, once q >= zs, output
// plane q from first-sweep planes q-1..q+1.  First-sweep values outside the
// grid's interior are 0, exactly what the first sweep's output holds there
// (zero Dirichlet), so two passes of this body equal four single sweeps bit for
// bit.  Columns: 2*HL halo lanes (one lane of V >= 2 columns, two of V = 1).
// SI / SO: storage of src / dst, as for Stencil3DHL; the first sweep's planes
// stay in registers in the accumulation type.
template <typename SI, int SHAPE, int V, int NY, int PF, typename SO = SI>
struct Stencil3DHL2 {
    using A = typename Storage<SI>::acc;
    static_assert(std::is_same<A, typename Storage<SO>::acc>::value, "one accumulation type");
    static_assert(PF >= 1, "prefetch at least one plane");
    static constexpr int R = 1;
    static constexpr int NT = ntaps3(SHAPE, R);
    static constexpr int HL = V >= 2 ? 1 : 2;
    static constexpr int OW = (kWave - 2 * HL) * V;
    static constexpr int NI = NY + 4;            // input rows y0-2 .. y0+NY+1
    static constexpr int NM = NY + 2;            // first-sweep rows y0-1 .. y0+NY
    static constexpr int NPI = 3 + PF;           // input planes in the ring
    static constexpr int PERIOD = clcm(NPI, 3);  // ticks until both rings realign
    static constexpr uint32_t ESI = sizeof(SI), ESO = sizeof(SO);

    struct PlaneI { A v[NI][V]; };
    struct PlaneM { A v[NM][V]; };

    struct Ctx {
        BufRsrc src, dst;
        int ny, nz, zs;
        uint32_t syi, szi, syo, szo;  // byte strides of src and dst
        int y0;
        int ze;
        bool single;          // one sweep only (first-sweep planes are the output)
        uint32_t lo;
        uint32_t so[V];
        bool xin[V];
    };

    template <bool VEC>
    __device__ static __forceinline__ void store_row(const Ctx& c, uint32_t roff, const A (&out)[V]) {
        buf_store_elems<SO, V, VEC>(c.dst, c.so, roff, out);
    }

    template <bool VEC>
    __device__ static __forceinline__ void load_plane(const Ctx& c, int z, PlaneI& P) {
        const uint32_t pz = static_cast<uint32_t>(z) * c.szi + static_cast<uint32_t>(c.y0 - 2) * c.syi;
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

    // tick KT of a period starting at q = zq0 (zq0 - (zs - 2) is a multiple of
    // PERIOD): input plane zq0+k sits in slot k % NPI, first-sweep plane zq0+k in
    // slot k % 3.
    template <bool VEC, int KT>
    __device__ static __forceinline__ void tick(const Ctx& c, const A* __restrict__ w, int zq0,
                                                PlaneI (&in)[NPI], PlaneM (&mid)[3]) {
        const int q = zq0 + KT;
        load_plane<VEC>(c, q + 2 + PF, in[(KT + 2 + PF) % NPI]);
        int zero = 0;
        asm volatile("" : "+s"(zero));
        const A* __restrict__ tw = w + zero;
        // first sweep: plane m = q+1, rows y0-1 .. y0+NY
        {
            const int m = q + 1;
            const bool zin = m >= 1 && m < c.nz - 1;
            PlaneM& M = mid[(KT + 1) % 3];
#pragma unroll
            for (int i = 0; i < NM; ++i) {
                const int y = c.y0 - 1 + i;
                const bool yin = zin && y >= 1 && y < c.ny - 1;
#pragma unroll
                for (int j = 0; j < V; ++j) {
                    A acc = 0;
#pragma unroll
                    for (int t = 0; t < NT; ++t) {
                        const Tap3 tp = tap3(SHAPE, t);
                        const PlaneI& P = in[(KT + 1 + tp.dz) % NPI];
                        acc = fmaa(tw[t], at(P.v, i + 1 + tp.dy, j, tp.dx), acc);
                    }
                    M.v[i][j] = (yin && c.xin[j]) ? acc : A(0);
                }
            }
            if (c.single && m >= c.zs && m < c.ze) {
                const uint32_t mz = static_cast<uint32_t>(m) * c.szo;
#pragma unroll
                for (int i = 1; i <= NY; ++i) {
                    const int y = c.y0 - 1 + i;
                    if (y < c.ny) store_row<VEC>(c, mz + static_cast<uint32_t>(y) * c.syo, M.v[i]);
                }
            }
        }
        // second sweep: output plane q, rows y0 .. y0+NY-1
        if (!c.single && q >= c.zs) {
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
                        const PlaneM& M = mid[(KT + 3 + tp.dz) % 3];
                        acc = fmaa(tw[t], at(M.v, i + 1 + tp.dy, j, tp.dx), acc);
                    }
                    out[j] = (yin && c.xin[j]) ? acc : A(0);
                }
                if (y < c.ny) store_row<VEC>(c, qz + static_cast<uint32_t>(y) * c.syo, out);
            }
        }
    }

    template <bool VEC, int KT>
    __device__ static __forceinline__ void period(const Ctx& c, const A* __restrict__ w, int zq0,
                                                  PlaneI (&in)[NPI], PlaneM (&mid)[3]) {
        if constexpr (KT < PERIOD) {
            tick<VEC, KT>(c, w, zq0, in, mid);
            period<VEC, KT + 1>(c, w, zq0, in, mid);
        }
    }

    template <bool VEC, int KT>
    __device__ static __forceinline__ void tail(const Ctx& c, const A* __restrict__ w, int zq0, int zend,
                                                PlaneI (&in)[NPI], PlaneM (&mid)[3]) {
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
        PlaneM mid[3];
        // ticks run q = zs-2 .. ze-1; the first tick needs input planes
        // zs-2 .. zs+PF-1 in slots 0 .. PF+1
#pragma unroll
        for (int k = 0; k < 2 + PF; ++k) load_plane<VEC>(c, zs - 2 + k, in[k]);
        int zq0 = zs - 2;
#pragma nounroll
        for (; zq0 + PERIOD <= ze; zq0 += PERIOD) period<VEC, 0>(c, w, zq0, in, mid);
        tail<VEC, 0>(c, w, zq0, ze, in, mid);
    }

    // One tile: strip `strip`, y-block `yb`, output planes [zs, ze) of a dense
    // grid that fits grid3d_fits_buffer, two sweeps from src to dst (single: one).
    __device__ static __forceinline__ void run_tile(const SI* __restrict__ src, SO* __restrict__ dst,
                                                    const Grid3D& g, bool vec, const A* __restrict__ w,
                                                    int strip, int yb, int zs, int ze, bool single) {
        const int lane = threadIdx.x & (kWave - 1);
        const uint32_t n = static_cast<uint32_t>(g.nx) * static_cast<uint32_t>(g.ny) * static_cast<uint32_t>(g.nz);
        Ctx c;
        c.src = buf_rsrc(src, n * ESI);
        c.dst = buf_rsrc(dst, n * ESO);
        c.ny = g.ny;
        c.nz = g.nz;
        c.zs = zs;
        c.ze = ze;
        c.single = single;
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
