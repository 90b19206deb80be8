// stencil3d_hl.hpp -- radius-1 3D stencils for narrow grids (tens to a few
// hundred columns), the tile body of the batched 3D launch.
//
// Same register-cache scheme as stencil3d.hpp (x-strip per wave, NY rows per
// lane, z streamed through a ring of planes, x-neighbours over DPP) with one
// change of layout that suits narrow grids: the strip's outer lanes are halo
// lanes.  Lane l of strip s holds columns x0 = s*OW + (l-1)*V .. x0+V-1 with
// OW = 62*V; lanes 0 and 63 only feed their neighbours and store nothing.
// Every load address is clamped into the grid, so the hot loop has no edge
// branch and no extra halo load per row; stores are masked per lane.
//
// Output rule (zero Dirichlet, as stencil3d.hpp): interior points get the
// stencil, the outer layer of the grid gets 0.
#pragma once
#include <hip/hip_runtime.h>

#include "ssam_device.hpp"
#include "stencil3d.hpp"

namespace ssam {

struct Grid3D {
    int nx, ny, nz;
    long long sy, sz;  // element strides of y and z
};

template <typename S, int SHAPE, int V, int NY, int PF>
struct Stencil3DHL {
    using A = typename Storage<S>::acc;
    static constexpr int R = 1;
    static constexpr int NT = ntaps3(SHAPE, R);
    static constexpr int NROW = NY + 2;     // rows cached per plane
    static constexpr int NP = 3 + PF;       // planes in the ring
    static constexpr int OW = (kWave - 2) * V;

    struct Plane { A v[NROW][V]; };

    struct Ctx {
        const S* src;
        S* dst;
        Grid3D g;
        int x0, y0;
        uint32_t xoff;        // byte offset of the lane's (clamped) first column
        uint32_t xe[V];       // per-element clamped byte offsets (VEC = false)
        bool store_lane;
    };

    // VEC: rows are V-aligned (nx % V == 0, aligned base), one vector access per
    // row; otherwise V dword accesses.  Both are branch-free.
    template <bool VEC>
    __device__ static __forceinline__ void load_plane(const Ctx& c, int z, Plane& P) {
        const int zz = clampi(z, 0, c.g.nz - 1);
        const char* base = reinterpret_cast<const char*>(c.src) +
                           static_cast<long long>(zz) * c.g.sz * static_cast<long long>(sizeof(S));
#pragma unroll
        for (int i = 0; i < NROW; ++i) {
            const int y = clampi(c.y0 - 1 + i, 0, c.g.ny - 1);
            const char* row = base + static_cast<long long>(y) * c.g.sy * static_cast<long long>(sizeof(S));
            if constexpr (VEC) {
                load_strip_vec<S, V>(reinterpret_cast<const S*>(row + c.xoff), P.v[i]);
            } else {
#pragma unroll
                for (int j = 0; j < V; ++j)
                    P.v[i][j] = Storage<S>::load(gld<S>(row + c.xe[j]));
            }
        }
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
        const bool zin = q >= 1 && q < c.g.nz - 1;
#pragma unroll
        for (int i = 0; i < NY; ++i) {
            const int y = c.y0 + i;
            const bool yin = zin && y >= 1 && y < c.g.ny - 1;
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
                const int x = c.x0 + j;
                out[j] = (yin && x >= 1 && x < c.g.nx - 1) ? acc : A(0);
            }
            if (c.store_lane && y < c.g.ny && q < c.g.nz) {
                char* row = reinterpret_cast<char*>(c.dst) +
                            (static_cast<long long>(q) * c.g.sz + static_cast<long long>(y) * c.g.sy) *
                                static_cast<long long>(sizeof(S));
                if constexpr (VEC) {
                    store_strip_vec<S, V>(reinterpret_cast<S*>(row + c.xoff), out);
                } else {
#pragma unroll
                    for (int j = 0; j < V; ++j)
                        if (c.x0 + j < c.g.nx)
                            gst<S>(row + static_cast<uint32_t>((c.x0 + j) * sizeof(S)),
                                Storage<S>::store(out[j]));
                }
            }
        }
    }

    template <bool VEC, int KT>
    __device__ static __forceinline__ void period(const Ctx& c, const A* __restrict__ w, int zq0,
                                                  int zend, Plane (&ring)[NP]) {
        if constexpr (KT < NP) {
            if (zq0 + KT < zend) tick<VEC, KT>(c, w, zq0, ring);
            period<VEC, KT + 1>(c, w, zq0, zend, ring);
        }
    }

    template <bool VEC>
    __device__ static __forceinline__ void stream(const Ctx& c, const A* __restrict__ w, int zs, int ze) {
        Plane ring[NP];
        // planes zs-1 .. zs+PF into the slots of the first period
#pragma unroll
        for (int k = 0; k < 2 + PF; ++k) load_plane<VEC>(c, zs - 1 + k, ring[k]);
        for (int zq0 = zs; zq0 < ze; zq0 += NP) period<VEC, 0>(c, w, zq0, ze, ring);
    }

    // One tile: strip `strip`, y-block `yb`, output planes [zs, ze) (wave-uniform).
    // vec: nx % V == 0 and src/dst aligned to V elements.
    __device__ static __forceinline__ void run_tile(const S* __restrict__ src, S* __restrict__ dst,
                                                    const Grid3D& g, bool vec, const A* __restrict__ w,
                                                    int strip, int yb, int zs, int ze) {
        const int lane = threadIdx.x & (kWave - 1);
        Ctx c;
        c.src = src;
        c.dst = dst;
        c.g = g;
        c.x0 = strip * OW + (lane - 1) * V;
        c.y0 = yb * NY;
        c.store_lane = lane >= 1 && lane < kWave - 1 && c.x0 < g.nx;
        // a lane whose strip leaves the grid reads the last aligned strip (its
        // values reach only lanes that store nothing or boundary columns)
        const int xv = max(min(c.x0, g.nx - V), 0);
        c.xoff = static_cast<uint32_t>(xv) * static_cast<uint32_t>(sizeof(S));
#pragma unroll
        for (int j = 0; j < V; ++j)
            c.xe[j] = static_cast<uint32_t>(clampi(c.x0 + j, 0, g.nx - 1)) * static_cast<uint32_t>(sizeof(S));
        if (V == 1 || vec) stream<true>(c, w, zs, ze);
        else stream<false>(c, w, zs, ze);
    }
};

}  // namespace ssam
