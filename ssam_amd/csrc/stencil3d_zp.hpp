// stencil3d_zp.hpp -- the z-partial-sum form of the 3D star stencils (radius 2:
// the 13-point kernel3d_3d13pt, ssai-3dstencil/tmp3d13pt.cu:74-230), after the
// reference's z partial-sum variants (tmp3d7pt-v1.cu:350-501, tmp3d7pt-v3.cu:62-228)
// recast for a wave64 register ring.
//
// The plane-ring kernel (stencil3d.hpp) keeps 2r+1+PF whole planes (NY+2r rows
// and their x-halos) in registers, since every plane is the centre of one output
// plane and a z-neighbour of 2r others.  For radius 2 that is 203 VGPRs, two waves
// per SIMD, 0.76 of the stream bandwidth.  Here each input plane is used once, at
// the tick it arrives: it adds its in-plane taps to its own output plane and its
// z taps to the 2r neighbouring output planes, whose partial sums stay open in a
// ring of 2r+1 accumulator planes (NY x V each).  Registers: 1+PF input planes and
// 2r+1 accumulator planes, which leaves room for twice the rows per lane (NY=4:
// fewer y-halo re-reads per output row) at the same occupancy.  Same taps, same
// zero boundary, same outputs as the plane ring up to the order of the FP
// additions; no residual (requests with one take the plane-ring kernel).
#pragma once
#include "stencil3d.hpp"

namespace ssam {

constexpr int zp_gcd(int a, int b) { return b == 0 ? a : zp_gcd(b, a % b); }

template <typename S, int R, int V, int NY, int PF, typename SO = S>
struct Stencil3DZP {
    using B = Stencil3D<S, k3Star, R, V, NY, PF, SO>;
    using A = typename B::A;
    using Plane = typename B::Plane;
    using Ctx = typename B::Ctx;
    static constexpr int NT = ntaps3(k3Star, R);
    static constexpr int NPL = 1 + PF;    // input planes held: the one consumed + PF in flight
    static constexpr int NO = 2 * R + 1;  // output planes with open partial sums
    static constexpr int PERIOD = NPL * NO / zp_gcd(NPL, NO);

    // descriptor index of the star tap at z offset dz (include/ssam.h order)
    static constexpr int tz(int dz) { return dz < 0 ? 6 * (-dz - 1) + 5 : 6 * (dz - 1) + 6; }
    // accumulator slot of output plane q + d at tick KT of a period (static)
    static constexpr int slot(int KT, int d) { return (KT + d + NO * (R + 1)) % NO; }

    template <bool CHK>
    __device__ static __forceinline__ void store_plane(const Ctx& c, int o, const A (&acc)[NY][V]) {
        const bool zin = o >= R && o < c.p.nz - R;
#pragma unroll
        for (int i = 0; i < NY; ++i) {
            const int y = c.y0 + i;
            if (y >= c.p.ny) continue;
            const bool yin = y >= R && y < c.p.ny - R;
            A out[V];
#pragma unroll
            for (int j = 0; j < V; ++j) {
                const int x = c.x0 + j;
                out[j] = zin && yin && x >= R && x < c.p.nx - R ? acc[i][j] : A(0);
            }
            char* row = reinterpret_cast<char*>(c.dst) +
                        (static_cast<long long>(o) * c.p.sz + static_cast<long long>(y) * c.p.sy) *
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

    // Tick KT consumes input plane q = q0 + KT: z taps into the open output planes
    // q-R .. q+R (q+R starts here), in-plane taps into q; output q-R is complete.
    template <bool CHK, int KT>
    __device__ static __forceinline__ void tick(const Ctx& c, const A* __restrict__ w, int q0, int zs,
                                                Plane (&ring)[NPL], A (&acc)[NO][NY][V]) {
        const int q = q0 + KT;
        B::template load_plane<CHK>(c, q + PF, ring[(KT + PF) % NPL]);
        int zero = 0;
        asm volatile("" : "+s"(zero));
        const A* __restrict__ tw = w + zero;
        const Plane& P = ring[KT % NPL];
#pragma unroll
        for (int i = 0; i < NY; ++i)
#pragma unroll
            for (int j = 0; j < V; ++j) {
                const A v = P.v[i + R][j];
#pragma unroll
                for (int d = -R; d <= R; ++d) {
                    if (d == 0) continue;
                    A& a = acc[slot(KT, d)][i][j];
                    // plane q is at z offset -d from output plane q + d
                    a = d == R ? tw[tz(-d)] * v : fmaa(tw[tz(-d)], v, a);
                }
                A& a0 = acc[slot(KT, 0)][i][j];
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const Tap3 tp = tap3(k3Star, t);
                    if (tp.dz != 0) continue;
                    a0 = fmaa(tw[t], B::at(c, P, i + R + tp.dy, j, tp.dx), a0);
                }
            }
        const int o = q - R;
        if (o >= zs) store_plane<CHK>(c, o, acc[slot(KT, -R)]);
    }

    template <bool CHK, int KT>
    __device__ static __forceinline__ void period(const Ctx& c, const A* __restrict__ w, int q0, int qend,
                                                  int zs, Plane (&ring)[NPL], A (&acc)[NO][NY][V]) {
        if constexpr (KT < PERIOD) {
            if (q0 + KT < qend) tick<CHK, KT>(c, w, q0, zs, ring, acc);
            period<CHK, KT + 1>(c, w, q0, qend, zs, ring, acc);
        }
    }

    // output planes [zs, ze): ticks consume input planes zs-R .. ze+R-1
    template <bool CHK>
    __device__ static __forceinline__ void stream(const Ctx& c, const A* __restrict__ w, int zs, int ze) {
        Plane ring[NPL];
        A acc[NO][NY][V];
#pragma unroll
        for (int s = 0; s < NO; ++s)
#pragma unroll
            for (int i = 0; i < NY; ++i)
#pragma unroll
                for (int j = 0; j < V; ++j) acc[s][i][j] = 0;
#pragma unroll
        for (int k = 0; k < PF; ++k) B::template load_plane<CHK>(c, zs - R + k, ring[k]);
        const int qend = ze + R;
        for (int q0 = zs - R; q0 < qend; q0 += PERIOD) period<CHK, 0>(c, w, q0, qend, zs, ring, acc);
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
            int hx = c.x0;
            if (lane == 0) hx = c.x0 - R;
            if (lane == kWave - 1) hx = c.x0 + V;
            c.hx = clampi(hx, 0, p.nx - 1);
        }
        const bool interior = p.vec_ok && (strip + 1) * kWave * V <= p.nx;
        if (interior)
            stream<false>(c, w, zs, ze);
        else
            stream<true>(c, w, zs, ze);
    }
};

template <typename S, int R, int V, int NY, int PF, typename SO = S>
__global__ __launch_bounds__(kWave * kWaves3D) void ssam3d_zp_kernel(
    Taps3<typename Storage<S>::acc, ntaps3(k3Star, R)> tp, const S* __restrict__ src, SO* __restrict__ dst,
    Params3D p) {
    (void)tp;
    using A = typename Storage<S>::acc;
    const A* __restrict__ w = (const A*)(__builtin_amdgcn_kernarg_segment_ptr());
    Stencil3DZP<S, R, V, NY, PF, SO>::run(src, dst, p, w);
}

}  // namespace ssam
