// stencil3d.hpp -- SSAM register-cache 3D stencils for gfx950 (wave64).
//
// Replaces kernel3d_3d7pt (ssai-3dstencil/tmp3d7pt-v2.cu:62-208), kernel3d_3d13pt
// (tmp3d13pt.cu:74-230) and kernel3d_poisson (tmpPoisson.cu:86-298), and runs this
// library's 19-point Jacobi operator with an FP64 residual.
//
// Execution model:
//  * a wave owns an x-strip of 64*V columns: lane l caches V consecutive x values
//    of NY+2r rows (y) of each plane; x-neighbours come from the adjacent lane via
//    DPP wave_shl/shr (lanes 0 and 63 take them from one extra load per row);
//  * the wave streams along z through a segment of planes, keeping a register ring
//    of 2r+1+PF planes (PF planes prefetched ahead), unrolled over the ring period
//    so ring slots are static registers;
//  * every output point is written once: interior points get the stencil, the
//    outer r layers of the grid get 0 (zero Dirichlet, as the reference gold's
//    untouched zero-initialised boundary, tmp3d7pt-v2.cu:284-296 + common.hpp:78-82);
//  * with a residual request each wave accumulates sum r^2 in FP64 and writes one
//    partial; a second kernel adds the partials in a fixed order.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "ssam_device.hpp"

namespace ssam {

enum Shape3D : int { k3Star = 0, k3FaceEdge = 1 };

// tap order of the descriptor (include/ssam.h)
//  STAR r: c, x-1, x+1, y-1, y+1, z-1, z+1, x-2, x+2, y-2, y+2, z-2, z+2
//  FACE_EDGE: c, 6 faces as STAR r=1, then (y-1,z-1) (y+1,z-1) (y-1,z+1) (y+1,z+1)
//             (x-1,z-1) (x+1,z-1) (x-1,z+1) (x+1,z+1) (x-1,y-1) (x+1,y-1) (x-1,y+1) (x+1,y+1)
struct Tap3 { int dz, dy, dx; };
__host__ __device__ constexpr Tap3 tap3(int shape, int i) {
    if (shape == k3FaceEdge && i >= 7) {
        const int e = i - 7;
        const int s0 = (e & 1) ? 1 : -1, s1 = (e & 2) ? 1 : -1;
        if (e < 4) return Tap3{s1, s0, 0};   // (y, z)
        if (e < 8) return Tap3{s1, 0, s0};   // (x, z)
        return Tap3{0, s1, s0};              // (x, y)
    }
    if (i == 0) return Tap3{0, 0, 0};
    const int k = (i - 1) / 6 + 1;           // distance
    const int a = ((i - 1) % 6) / 2;         // axis 0=x 1=y 2=z
    const int sgn = ((i - 1) % 2) ? 1 : -1;
    return a == 0 ? Tap3{0, 0, sgn * k} : (a == 1 ? Tap3{0, sgn * k, 0} : Tap3{sgn * k, 0, 0});
}
constexpr int ntaps3(int shape, int r) { return shape == k3FaceEdge ? 19 : 6 * r + 1; }

template <typename A, int NT> struct Taps3 { A w[NT]; };

struct Params3D {
    int nx, ny, nz;
    long long sy, sz;       // element strides of y and z
    int nstrips, nyb, ntiles;
    int zseg;               // output planes per tile
    int nblocks;
    int vec_ok;
    double* partials;       // per-tile residual partials (nullptr: none)
    double res_scale;       // r = res_scale * (out - centre)
};

constexpr int kWaves3D = 4;

// S: storage of src; SO: storage of dst (default S) -- a chain of sweeps over
// BF16 / FP16 grids keeps its running state in FP32 between launches.
template <typename S, int SHAPE, int R, int V, int NY, int PF, typename SO = S>
struct Stencil3D {
    using A = typename Storage<S>::acc;
    static_assert(std::is_same<A, typename Storage<SO>::acc>::value, "one accumulation type");
    static constexpr int NT = ntaps3(SHAPE, R);
    static constexpr int NROW = NY + 2 * R;   // rows cached per plane
    static constexpr int NP = 2 * R + 1 + PF; // planes in the ring

    struct Plane {
        A v[NROW][V];
        A halo[NROW][R];  // lane 0: x0-R..x0-1 ; lane 63: x0+V..x0+V+R-1
    };

    struct Ctx {
        const S* src;
        SO* dst;
        Params3D p;
        int lane, x0, y0;
        uint32_t xoff, xoff_out;  // byte offsets of the lane's columns in a src / dst row
        int hx;  // first column of this lane's halo load (always in [0, nx))
        bool full;
    };

    // CHK = false: every lane of the wave holds a full, aligned strip, so rows load
    // with one unconditional vector access (no per-lane branch in the hot loop).
    template <bool CHK>
    __device__ static __forceinline__ void load_plane(const Ctx& c, int z, Plane& P) {
        const int zz = clampi(z, 0, c.p.nz - 1);
        const char* base = reinterpret_cast<const char*>(c.src) +
                           static_cast<long long>(zz) * c.p.sz * static_cast<long long>(sizeof(S));
#pragma unroll
        for (int i = 0; i < NROW; ++i) {
            const int y = clampi(c.y0 - R + i, 0, c.p.ny - 1);
            const char* row = base + static_cast<long long>(y) * c.p.sy * static_cast<long long>(sizeof(S));
            if (!CHK || (c.full && c.p.vec_ok)) {
                load_strip_vec<S, V>(reinterpret_cast<const S*>(row + c.xoff), P.v[i]);
            } else {
#pragma unroll
                for (int j = 0; j < V; ++j) {
                    const uint32_t x = static_cast<uint32_t>(clampi(c.x0 + j, 0, c.p.nx - 1));
                    P.v[i][j] = Storage<S>::load(gld<S>(row + x * sizeof(S)));
                }
            }
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const uint32_t hx = static_cast<uint32_t>(min(c.hx + k, c.p.nx - 1));
                P.halo[i][k] = Storage<S>::load(gld<S>(row + hx * sizeof(S)));
            }
        }
    }

    // value at (row i of plane P, column j + dx) for dx in [-R, R]
    __device__ static __forceinline__ A at(const Ctx& c, const Plane& P, int i, int j, int dx) {
        const int cc = j + dx;
        if (cc >= 0 && cc < V) return P.v[i][cc];
        if (cc < 0) {
            const int k = -cc - 1;  // 0 = x0-1
            const A d = from_left(P.v[i][V - 1 - k]);
            return c.lane == 0 ? P.halo[i][R - 1 - k] : d;
        }
        const int k = cc - V;       // 0 = x0+V
        const A d = from_right(P.v[i][k]);
        return c.lane == kWave - 1 ? P.halo[i][k] : d;
    }

    template <bool CHK, int KT>
    __device__ static __forceinline__ void tick(const Ctx& c, const A* __restrict__ w, int zq0,
                                                Plane (&ring)[NP], double& res) {
        const int q = zq0 + KT;  // output plane
        // prefetch plane q + R + PF into the slot of plane q - R - 1
        load_plane<CHK>(c, q + R + PF, ring[(KT + 2 * R + PF) % NP]);
        int zero = 0;
        asm volatile("" : "+s"(zero));
        const A* __restrict__ tw = w + zero;
        const bool zin = q >= R && q < c.p.nz - R;
#pragma unroll
        for (int i = 0; i < NY; ++i) {
            const int y = c.y0 + i;
            const bool yin = y >= R && y < c.p.ny - R;
            A out[V];
#pragma unroll
            for (int j = 0; j < V; ++j) {
                A acc = 0;
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const Tap3 tp = tap3(SHAPE, t);
                    const Plane& P = ring[(KT + R + tp.dz + NP) % NP];
                    acc = fmaa(tw[t], at(c, P, i + R + tp.dy, j, tp.dx), acc);
                }
                const int x = c.x0 + j;
                const bool in = zin && yin && x >= R && x < c.p.nx - R;
                if (c.p.partials && in) {
                    const double r = c.p.res_scale *
                        (static_cast<double>(acc) - static_cast<double>(ring[(KT + R) % NP].v[i + R][j]));
                    res += r * r;
                }
                out[j] = in ? acc : A(0);
            }
            if (q < c.p.nz && y < c.p.ny && zq0 + KT >= 0) {
                char* row = reinterpret_cast<char*>(c.dst) +
                            (static_cast<long long>(q) * c.p.sz + static_cast<long long>(y) * c.p.sy) *
                                static_cast<long long>(sizeof(SO));
                if (!CHK || (c.full && c.p.vec_ok)) {
                    store_strip_vec<SO, V>(reinterpret_cast<SO*>(row + c.xoff_out), out);
                } else {
#pragma unroll
                    for (int j = 0; j < V; ++j)
                        if (c.x0 + j < c.p.nx && c.x0 + j >= 0)
                            gst<SO>(row + c.xoff_out + static_cast<uint32_t>(j * sizeof(SO)),
                                    Storage<SO>::store(out[j]));
                }
            }
        }
    }

    template <bool CHK, int KT>
    __device__ static __forceinline__ void period(const Ctx& c, const A* __restrict__ w, int zq0,
                                                  int zend, Plane (&ring)[NP], double& res) {
        if constexpr (KT < NP) {
            if (zq0 + KT < zend) tick<CHK, KT>(c, w, zq0, ring, res);
            period<CHK, KT + 1>(c, w, zq0, zend, ring, res);
        }
    }

    template <bool CHK>
    __device__ static __forceinline__ double stream(const Ctx& c, const A* __restrict__ w, int zs,
                                                    int ze) {
        Plane ring[NP];
        // planes zs-R .. zs+R+PF-1 into the slots of the first period
#pragma unroll
        for (int k = 0; k < 2 * R + PF; ++k) load_plane<CHK>(c, zs - R + k, ring[k]);
        double res = 0.0;
        for (int zq0 = zs; zq0 < ze; zq0 += NP) period<CHK, 0>(c, w, zq0, ze, ring, res);
        return res;
    }

    // One tile: x-strip `strip`, y-block `yb`, output planes [zs, ze) of the grid
    // described by p (all wave-uniform).  Returns this lane's residual^2 sum.
    __device__ static __forceinline__ double run_tile(const S* __restrict__ src, SO* __restrict__ dst,
                                                      const Params3D& p, const A* __restrict__ w,
                                                      int strip, int yb, int zs, int ze) {
        const int lane = threadIdx.x & (kWave - 1);
        Ctx c;
        c.src = src;
        c.dst = dst;
        c.p = p;
        c.lane = lane;
        c.x0 = strip * kWave * V + lane * V;
        c.y0 = yb * NY;
        c.full = c.x0 + V <= p.nx;
        c.xoff = static_cast<uint32_t>(c.x0) * static_cast<uint32_t>(sizeof(S));
        c.xoff_out = static_cast<uint32_t>(c.x0) * static_cast<uint32_t>(sizeof(SO));
        {
            int hx = c.x0;  // other lanes: any in-range column (their halo is unused)
            if (lane == 0) hx = c.x0 - R;
            if (lane == kWave - 1) hx = c.x0 + V;
            c.hx = clampi(hx, 0, p.nx - 1);
        }
        // the strip is full for every lane when its last column is inside the grid
        const bool interior = p.vec_ok && (strip + 1) * kWave * V <= p.nx;
        return interior ? stream<false>(c, w, zs, ze) : stream<true>(c, w, zs, ze);
    }

    __device__ static void run(const S* __restrict__ src, SO* __restrict__ dst, const Params3D& p,
                               const A* __restrict__ w) {
        const int lane = threadIdx.x & (kWave - 1);
        const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
        int blk = blockIdx.x;
        {
            const int n8 = (p.nblocks / 8) * 8;
            if (blk < n8) blk = (blk % 8) * (n8 / 8) + blk / 8;
        }
        const int tile = blk * kWaves3D + wv;
        if (tile >= p.ntiles) return;
        const int strip = tile % p.nstrips;
        const int rest = tile / p.nstrips;
        const int yb = rest % p.nyb;
        const int zs = (rest / p.nyb) * p.zseg;
        const int ze = min(p.nz, zs + p.zseg);
        double res = run_tile(src, dst, p, w, strip, yb, zs, ze);
        if (p.partials) {
            // wave reduction in a fixed order, then one partial per tile
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) res += __shfl_xor(res, off);
            if (lane == 0) p.partials[tile] = res;
        }
    }
};

template <typename S, int SHAPE, int R, int V, int NY, int PF, typename SO = S>
__global__ __launch_bounds__(kWave * kWaves3D) void ssam3d_kernel(
    Taps3<typename Storage<S>::acc, ntaps3(SHAPE, R)> tp, const S* __restrict__ src,
    SO* __restrict__ dst, Params3D p) {
    (void)tp;
    using A = typename Storage<S>::acc;
    const A* __restrict__ w = (const A*)(__builtin_amdgcn_kernarg_segment_ptr());
    Stencil3D<S, SHAPE, R, V, NY, PF, SO>::run(src, dst, p, w);
}

}  // namespace ssam
