// batch_common.hpp -- declarations shared by the batched launches (batch.hip:
// plan and host side; batch2d.hip / batch3d.hip: the persistent and grid-mode
// kernels): the 2D launch kinds, the item records, the XCD work queues and the
// 3D tile shapes.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "ssam_internal.hpp"
#include "stencil2d_dispatch.hpp"
#include "stencil3d.hpp"
#include "stencil3d_hl.hpp"

namespace ssam {

// ------------------------------------------------------------------ 2D menu
// Shapes the persistent 2D kernel runs in one launch (mixes of 5-point stars and
// 3x3 / 5x5 / 11x11 boxes); other shapes fall back to their own launch.
// k2Box2M / k2Box5M: the 5x5 / 11x11 box on the matrix cores (BF16/FP16 storage).
enum Kind2D : int { k2Star1 = 0, k2Box1 = 1, k2Box2 = 2, k2Box5 = 3, k2Box5M = 4, k2Box2M = 5, k2None = -1 };
constexpr int kNumKind2D = 4;   // tile-height slots (k2Box2M / k2Box5M use k2Box2's / k2Box5's)
constexpr int kNumLaunch2D = 6;  // launch kinds of the split form (one persistent launch each)

inline int seg_slot(int kind) { return kind == k2Box5M ? k2Box5 : kind == k2Box2M ? k2Box2 : kind; }

inline int kind2d(int shape, int r) {
    if (shape == kStar && r == 1) return k2Star1;
    if (shape == kBox && r == 1) return k2Box1;
    if (shape == kBox && r == 2) return k2Box2;
    if (shape == kBox && r == 5) return k2Box5;
    return k2None;
}

// BF16/FP16 box tiles take 4 columns per lane (248-column strips, as FP32): the
// batch's grids are narrow enough that 496-column strips leave lanes idle
// (measured 6% faster on the whole uneven batch, 30% on its 5x5 boxes).  The
// 5-point star keeps 8 (one 16-byte row access per lane): its BF16 batch items
// ran 4% faster that way once edge tiles stopped waiting at every row.
template <typename S, int R, int SHAPE> struct Inst2D {
    static constexpr int V =
        sizeof(S) == 2 && !(SHAPE == kStar && R == 1) ? Lanes2D<S>::VN : Lanes2D<S>::V;
    using UR = Unroll2D<R, SHAPE, V, 1>;
    using K = Stencil2D<S, R, SHAPE, V, 1, UR::PF, UR::RING, UR::U>;
    using G = Geom2D<R, V, 1>;
};

struct BatchItem2D {
    const void* src;
    void* dst;
    Params2D p;
    ParamsM pm;    // k2Box5M tiles
    int kind;
    int taps_off;  // offset into the taps array, in accumulation-type elements
};

struct BatchItem3D {
    const void* src;
    void* dst;
    void* work;
    Grid3D g;
    int nstrips, nyb;
    int vec;       // nx % V == 0 and every buffer aligned to V elements
    int steps;     // passes over the grid (sweeps / sweeps per pass, rounded up)
    int taps_off;
    int var;       // tile shape of this item's tiles (a kVar3D index)
    int fp32_state;  // BF16 / F16 item of several passes: the state between passes
                     // ping-pongs between two FP32 grids in `work`
};

constexpr int kBatchWavesPerBlock = 4;
constexpr int kQueues = 8;  // one per XCD

// First tile index of each queue (queue q holds tiles [beg[q], beg[q+1])).
// affinity 1: a wave starts at its own XCD's queue; 0: every wave starts at
// queue 0, so the whole device works through the tiles in table order.
// alternate (3D): odd steps walk the table back to front.
struct QueueSet {
    int beg[kQueues + 1];
    int affinity;
    int alternate;
};

// Physical XCD of this wave (hardware register XCC_ID, bits 3:0).
__device__ __forceinline__ int xcd_id() {
    return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & (kQueues - 1);
}

// Each queue's counter sits on a 256-byte line of its own: atomics on one
// address serialise where they execute, atomics on different lines do not.
constexpr int kCounterStride = 64;                      // ints
constexpr int kCounterSet = kQueues * kCounterStride;   // ints per launch

// Lane 0 claims the next tile of a queue (one atomic); other lanes return len.
__device__ __forceinline__ int claim(int* counter, int len) {
    int t = len;
    if ((threadIdx.x & (kWave - 1)) == 0) t = atomicAdd(counter, 1);
    return t;
}

// The same for a queue that is likely drained already (another XCD's): one
// relaxed load first, so finding it empty costs no atomic.
__device__ __forceinline__ int claim_checked(int* counter, int len) {
    int t = len;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        if (__hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < len)
            t = atomicAdd(counter, 1);
    }
    return t;
}

// Run body(tile) for tiles of this XCD's queue, then of the others in turn,
// until every queue is empty.  counters: kCounterSet ints, zero at launch.
// reverse: walk queues and tiles back to front (the 3D sweeps alternate, so a
// step starts on the grids the previous step wrote last, still in the last-level
// cache).  AHEAD: the claim of a wave's next tile is issued before its current
// tile runs, so the atomic's round trip overlaps the tile (for short tiles; with
// few long tiles per wave, a tile held in reserve by a busy wave lengthens the
// tail).
template <bool AHEAD, typename F>
__device__ __forceinline__ void drain_queues(const QueueSet& qs, int* __restrict__ counters, bool reverse,
                                             F&& body) {
    const int x = qs.affinity ? xcd_id() : 0;
    // one flat loop (a nested queue loop around the tile body costs the tile
    // registers): k counts the queues already drained
    int k = 0;
    int q = reverse ? kQueues - 1 - x : x;
    int beg = qs.beg[q], len = qs.beg[q + 1] - beg;
    int n = __builtin_amdgcn_readfirstlane(claim(counters + q * kCounterStride, len));
    for (;;) {
        if (n >= len) {
            if (++k == kQueues) break;
            const int qq = (x + k) & (kQueues - 1);
            q = reverse ? kQueues - 1 - qq : qq;
            beg = qs.beg[q];
            len = qs.beg[q + 1] - beg;
            n = __builtin_amdgcn_readfirstlane(claim_checked(counters + q * kCounterStride, len));
            continue;
        }
        if constexpr (AHEAD) {
            const int next = claim(counters + q * kCounterStride, len);
            body(reverse ? beg + len - 1 - n : beg + n);
            n = __builtin_amdgcn_readfirstlane(next);
        } else {
            body(reverse ? beg + len - 1 - n : beg + n);
            n = __builtin_amdgcn_readfirstlane(claim(counters + q * kCounterStride, len));
        }
    }
}

template <typename S> constexpr bool kHalf2D = std::is_same<S, bf16_t>::value || std::is_same<S, f16_t>::value;

// Batched 3D tile shapes: (columns per lane, rows per lane, planes prefetched).
// Narrow grids favour few columns per lane (62*V-column strips waste little of a
// 64..256-column row); more rows per lane cut the y-halo re-reads.
// wpe: waves per SIMD the register allocation must leave room for (bounds the
// compiler's load hoisting to what each shape needs).
// t: sweeps per pass (2: temporal blocking, Stencil3DHL2).
// s: the sweeps a kernel of a three-sweep shape runs (0: t); the shape's passes
// of two and one sweep (the tail of a count) launch kernels of their own on the
// same tiles (one kernel running all three tile bodies spilled registers).
struct Var3D { int v, ny, pf, wpe, t; int s = 0; };
// Variant index kVarPair3D: per item, the better of shapes kPairA (narrow
// grids) and kPairB (wider ones), in one kernel (pair_choice below); its table
// entry repeats kPairA's shape.  The shapes kept are those that measured within
// reach of the best on MI355X (a wider menu only lengthened the build).
constexpr int kVarPair3D = 4;
constexpr int kPairA = 2, kPairB = 3;
// kVarT2: two sweeps per pass, 2 columns x 6 rows per lane (always an ordinary
// grid).  kVarT3: three sweeps per pass for 7-point items (Stencil3DHLT), 2
// columns x 4 rows per lane; its passes of two or one sweep (the tail of a
// count) run the same tiles; 19-point items take kVarT2 instead.
constexpr int kVarT2 = 5;
constexpr int kVarT3 = 6;
constexpr int kVarT3S2 = 7, kVarT3S1 = 8;  // kVarT3's passes of two / one sweep (launch only)
constexpr Var3D kVar3D[] = {{1, 8, 2, 4, 1}, {2, 8, 2, 2, 1}, {2, 12, 2, 2, 1}, {4, 8, 1, 2, 1},
                            {2, 12, 2, 2, 1}, {2, 6, 1, 3, 2},  {2, 4, 1, 2, 3},  {2, 4, 1, 2, 3, 2},
                            {2, 4, 1, 2, 3, 1}};
constexpr int kNumVar3D = 7;  // shapes a plan may take (the tuner's and SSAM_BATCH3D_VARIANT's range)
constexpr int kDefaultVar3D = kVarT2;
constexpr int kDefaultZseg3D = 16;
constexpr int kDefaultSeg2D = 64;

// (by the accumulation type: a BF16 / F16 chain and its FP32 state share a shape)
template <typename S, int VAR> struct Var3DT {
    using A = typename Storage<S>::acc;
    // FP64 runs one column per lane (a plane row of NY+2 doubles per lane)
    static constexpr int V = sizeof(A) == 8 ? 1 : kVar3D[VAR].v;
    static constexpr int NY = kVar3D[VAR].ny, PF = kVar3D[VAR].pf;
    static constexpr int WPE = sizeof(A) == 8 ? (NY >= 16 || kVar3D[VAR].t >= 2 ? 1 : 2) : kVar3D[VAR].wpe;
    static constexpr int T = kVar3D[VAR].t;
    static constexpr int NSW = kVar3D[VAR].s ? kVar3D[VAR].s : T;  // sweeps this kernel runs
};

// Sweeps in each pass of a 3D item of `steps` sweeps with a variant of t sweeps
// per pass: t = 2 -- twos, the last pass of an odd count one; t = 3 -- threes,
// four as 2+2, the rest twos and ones (as ssam_stencil3d plans its launches).
inline int pass_sweeps3d(int steps, int variant, int pass) {
    const int t = kVar3D[variant].t;
    int left = steps;
    for (int p = 0;; ++p) {
        const int n = t == 3 ? ((left >= 3 && left != 4) ? 3 : (left >= 2 ? 2 : 1))
                             : (t == 2 && left >= 2 ? 2 : 1);
        if (p == pass || left <= n) return p == pass ? n : 0;
        left -= n;
    }
}
inline int passes3d(int steps, int variant) {
    int n = 0;
    for (int left = steps; left > 0; left -= pass_sweeps3d(left, variant, 0)) ++n;
    return n;
}
// Halo lanes per strip side of a variant (columns of every sweep of a pass).
inline int var3d_hl(int variant, int v) {
    const int t = kVar3D[variant].t;
    return t == 3 ? (3 + v - 1) / v : (t == 2 && v == 1 ? 2 : 1);
}

// Columns per lane of a variant for storage of `elem_bytes` bytes.
inline int var3d_v(int variant, int elem_bytes) { return elem_bytes == 8 ? 1 : kVar3D[variant].v; }

// The pair's shape for a grid nx columns wide: the one with the least estimated
// time per point, lane slots (62*V-column strips, the last one partly empty) x
// y-halo rows ((NY+2)/NY) / the measured rate of a full lane slot of that
// shape (wider V issues fewer instructions per byte: 512, 630 and 797 Gpoints/s
// per lane-slot share for V = 1, 2, 4 on one MI355X).
inline int pair_choice(long long nx, int elem_bytes) {
    auto cost = [&](int var) {
        const int V = var3d_v(var, elem_bytes), NY = kVar3D[var].ny;
        const double rate = V == 1 ? 512.0 : (V == 2 ? 630.0 : 797.0);
        const long long ow = (kWave - 2) * V;
        const double slots = static_cast<double>((nx + ow - 1) / ow * ow);
        return slots * (NY + 2) / NY / rate;
    };
    return cost(kPairB) < cost(kPairA) ? kPairB : kPairA;
}

// Kernel entry points by launch parameters (host stubs for hipLaunchKernel and
// the occupancy query).  storage: SSAM_F32 / F64 / BF16 / F16; kind: a Kind2D
// (split form) or -1 (fused); shape3d: 0 7-point star, 1 19-point face+edge;
// grid: one tile per wave of an ordinary grid instead of a persistent grid.
const void* batch2d_kernel_fn(int storage, int kind, bool grid);
// 3D passes read `storage` and write `out_storage` (the same, or a 2-byte type to
// / from its FP32 state; those pairs exist for the two-sweep shape only).
const void* batch3d_kernel_fn(int storage, int out_storage, int variant, int shape3d, bool grid);
const void* batch3d_t2_kernel_fn(int storage, int out_storage, int shape3d);
const void* batch3d_t2_kernel_fn_bf16(int storage, int out_storage, int shape3d);
const void* batch3d_t2_kernel_fn_f16(int storage, int out_storage, int shape3d);
// The kernel of a three-sweep shape's pass (variant kVarT3, kVarT3S2 or kVarT3S1;
// 7-point star) by storage pair (batch3d_t3*.hip).
const void* batch3d_t3_kernel_fn(int storage, int out_storage, int variant);
const void* batch3d_t3_kernel_fn_bf16(int storage, int out_storage, int variant);
const void* batch3d_t3_kernel_fn_f16(int storage, int out_storage, int variant);

}  // namespace ssam
