// batch.hip -- many grids of uneven sizes and mixed stencil shapes in one
// persistent launch fed by a tile work queue (2D), multi-step 3D sweeps in one
// persistent launch per step, and the whole sequence captured in a hipGraph.
//
// No counterpart in the reference (one grid, one launch per harness call).
//
//  * Plan: every 2D item is cut into tiles (column strip x segment of rows) of the
//    same geometry the single-grid launch uses; all tiles of one storage type go
//    into one table, ordered by estimated cost (longest first) so the queue
//    drains evenly.  3D items likewise (x-strip x y-block x z-segment).
//  * Kernel: a fixed grid of waves; each wave takes the next tile index from a
//    device counter (one vector atomic from lane 0), reads its item (scalar
//    loads), and runs the same register-cache stream as the single-grid kernel for
//    that item's shape; a wave leaves when the counter passes the table.
//  * Graph: counter reset (a fill kernel, no memset), the 2D launches, the per-step
//    3D launches and any per-item fallback launches are captured once on a private
//    stream and replayed on the caller's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "ssam_internal.hpp"
#include "stencil2d_dispatch.hpp"
#include "stencil3d.hpp"

namespace ssam {

// ------------------------------------------------------------------ 2D menu
// Shapes the persistent 2D kernel runs in one launch (BASELINE-style mixes of
// 5-point stars and 3x3 / 5x5 / 11x11 boxes); other shapes fall back to their own
// launch.
enum Kind2D : int { k2Star1 = 0, k2Box1 = 1, k2Box2 = 2, k2Box5 = 3, k2None = -1 };

inline int kind2d(int shape, int r) {
    if (shape == kStar && r == 1) return k2Star1;
    if (shape == kBox && r == 1) return k2Box1;
    if (shape == kBox && r == 2) return k2Box2;
    if (shape == kBox && r == 5) return k2Box5;
    return k2None;
}

template <typename S, int R, int SHAPE> struct Inst2D {
    static constexpr int V = Lanes2D<S>::V;
    using UR = Unroll2D<R, SHAPE, V, 1>;
    using K = Stencil2D<S, R, SHAPE, V, 1, UR::PF, UR::RING, UR::U>;
    using G = Geom2D<R, V, 1>;
};

struct BatchItem2D {
    const void* src;
    void* dst;
    Params2D p;
    int kind;
    int taps_off;  // offset into the taps array, in accumulation-type elements
};

struct BatchItem3D {
    const void* src;
    void* dst;
    void* work;
    Params3D p;
    int kind;      // 0: 7-point star, 1: 19-point face+edge
    int steps;
    int taps_off;
};

constexpr int kBatchWavesPerBlock = 4;

__device__ __forceinline__ int next_tile(int* counter) {
    int t = 0;
    if ((threadIdx.x & (kWave - 1)) == 0) t = atomicAdd(counter, 1);
    return __builtin_amdgcn_readfirstlane(t);
}

template <typename S>
__global__ __launch_bounds__(kWave * kBatchWavesPerBlock) __attribute__((amdgpu_waves_per_eu(2)))
void batch2d_kernel(const BatchItem2D* __restrict__ items, const int4* __restrict__ tiles, int ntiles,
                    int* __restrict__ counter, const typename Storage<S>::acc* __restrict__ taps) {
    for (;;) {
        const int t = next_tile(counter);
        if (t >= ntiles) break;
        const int4 e = tiles[t];
        const int it = __builtin_amdgcn_readfirstlane(e.x);
        const int strip = __builtin_amdgcn_readfirstlane(e.y);
        const int seg = __builtin_amdgcn_readfirstlane(e.z);
        const BatchItem2D& b = items[it];
        const S* src = static_cast<const S*>(b.src);
        S* dst = static_cast<S*>(b.dst);
        const typename Storage<S>::acc* w = taps + b.taps_off;
        switch (b.kind) {
            case k2Star1: Inst2D<S, 1, kStar>::K::run_tile(src, dst, b.p, w, strip, seg); break;
            case k2Box1: Inst2D<S, 1, kBox>::K::run_tile(src, dst, b.p, w, strip, seg); break;
            case k2Box2: Inst2D<S, 2, kBox>::K::run_tile(src, dst, b.p, w, strip, seg); break;
            default: Inst2D<S, 5, kBox>::K::run_tile(src, dst, b.p, w, strip, seg); break;
        }
    }
}

// Batched 3D grids are narrow (tens to a few hundred columns): 2 columns per lane
// (128-column strips) keep most lanes busy where 4 (256 columns) would idle half.
template <typename S> struct Lanes3Db;
template <> struct Lanes3Db<float> { static constexpr int V = 2; };
template <> struct Lanes3Db<double> { static constexpr int V = 1; };
constexpr int kNY3 = 4;   // rows per lane (radius 1)
constexpr int kPF3 = 1;

template <typename S>
__global__ __launch_bounds__(kWave * kBatchWavesPerBlock) __attribute__((amdgpu_waves_per_eu(2)))
void batch3d_kernel(const BatchItem3D* __restrict__ items, const int4* __restrict__ tiles, int ntiles,
                    int* __restrict__ counter, const S* __restrict__ taps, int step) {
    constexpr int V = Lanes3Db<S>::V;
    for (;;) {
        const int t = next_tile(counter);
        if (t >= ntiles) break;
        const int4 e = tiles[t];
        const int it = __builtin_amdgcn_readfirstlane(e.x);
        const int strip = __builtin_amdgcn_readfirstlane(e.y);
        const int yb = __builtin_amdgcn_readfirstlane(e.z);
        const int zs = __builtin_amdgcn_readfirstlane(e.w);
        const BatchItem3D& b = items[it];
        // ping-pong so the last sweep writes dst (same rule as ssam_stencil3d)
        const void* in = step == 0 ? b.src : (((b.steps - step) % 2 == 0) ? b.dst : b.work);
        void* out = ((b.steps - 1 - step) % 2 == 0) ? b.dst : b.work;
        const int ze = min(b.p.nz, zs + b.p.zseg);
        const S* w = taps + b.taps_off;
        if (b.kind == 0)
            (void)Stencil3D<S, k3Star, 1, V, kNY3, kPF3>::run_tile(static_cast<const S*>(in),
                                                                   static_cast<S*>(out), b.p, w,
                                                                   strip, yb, zs, ze);
        else
            (void)Stencil3D<S, k3FaceEdge, 1, V, kNY3, kPF3>::run_tile(static_cast<const S*>(in),
                                                                       static_cast<S*>(out), b.p,
                                                                       w, strip, yb, zs, ze);
    }
}

__global__ void reset_counters_kernel(int* __restrict__ c, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c[i] = 0;
}

}  // namespace ssam

using namespace ssam;

// ------------------------------------------------------------------ plan
struct ssam_batch_plan_s {
    int device = 0;
    struct Group {  // one persistent launch: items of one storage type
        int ndim = 2;
        int storage = 0;
        void* d_items = nullptr;
        void* d_tiles = nullptr;
        void* d_taps = nullptr;
        int ntiles = 0;
        int steps = 1;  // 3D
    };
    std::vector<Group> groups;
    std::vector<ssam_batch_item> fallback;
    std::vector<std::vector<double>> fallback_taps;
    int ncounters = 0;
    void* d_counters = nullptr;
    int64_t tiles2d = 0, tiles3d = 0, point_updates = 0;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int persistent_blocks = 1024;
};

namespace {

int elem_size_b(int dtype) {
    switch (dtype) {
        case SSAM_F32: return 4;
        case SSAM_F64: return 8;
        case SSAM_BF16:
        case SSAM_F16: return 2;
        default: return 0;
    }
}

template <typename S, int R, int SHAPE>
void tiles_for_2d(const ssam_batch_item& it, int idx, int seg_rows, Params2D* p,
                  std::vector<std::pair<double, int4>>* out) {
    using I = Inst2D<S, R, SHAPE>;
    p->width = static_cast<int>(it.nx);
    p->height = static_cast<int>(it.ny);
    p->pitch = it.pitch;
    p->nstrips = static_cast<int>((it.nx + I::G::OW - 1) / I::G::OW);
    p->seg_rows = seg_rows;
    p->boundary = kZero;
    p->vec_ok = (it.pitch % I::V == 0) && (reinterpret_cast<uintptr_t>(it.src) % 16 == 0) &&
                (reinterpret_cast<uintptr_t>(it.dst) % 16 == 0);
    const int nsegs = (p->height + seg_rows - 1) / seg_rows;
    p->ntiles = p->nstrips * nsegs;
    p->nblocks = 0;
    const double taps = SHAPE == kStar ? 4 * R + 1 : (2 * R + 1) * (2 * R + 1);
    for (int sg = 0; sg < nsegs; ++sg) {
        const int rows = std::min(seg_rows, p->height - sg * seg_rows);
        for (int st = 0; st < p->nstrips; ++st)
            out->push_back({static_cast<double>(rows + 2 * R) * (taps + 8.0), int4{idx, st, sg, 0}});
    }
}

template <typename S>
int build_group2d(ssam_batch_plan_s* P, const std::vector<int>& ids, const ssam_batch_item* items,
                  int storage) {
    using A = typename Storage<S>::acc;
    std::vector<BatchItem2D> h(ids.size());
    std::vector<A> taps;
    std::vector<std::pair<double, int4>> tl;
    const int seg_rows = 32;
    for (size_t k = 0; k < ids.size(); ++k) {
        const ssam_batch_item& it = items[ids[k]];
        const int R = it.radius, K = 2 * R + 1;
        const int shape = it.shape == SSAM_SHAPE_STAR ? kStar : kBox;
        BatchItem2D& b = h[k];
        b.src = it.src;
        b.dst = it.dst;
        b.kind = kind2d(shape, R);
        b.taps_off = static_cast<int>(taps.size());
        for (int m = 0; m < K; ++m)
            for (int n = 0; n < K; ++n) {
                double v = it.coeffs[m * K + n];
                if (shape == kStar && m != R && n != R) v = 0.0;
                taps.push_back(static_cast<A>(v));
            }
        const int idx = static_cast<int>(k);
        switch (b.kind) {
            case k2Star1: tiles_for_2d<S, 1, kStar>(it, idx, seg_rows, &b.p, &tl); break;
            case k2Box1: tiles_for_2d<S, 1, kBox>(it, idx, seg_rows, &b.p, &tl); break;
            case k2Box2: tiles_for_2d<S, 2, kBox>(it, idx, seg_rows, &b.p, &tl); break;
            default: tiles_for_2d<S, 5, kBox>(it, idx, seg_rows, &b.p, &tl); break;
        }
        P->point_updates += it.nx * it.ny;
    }
    // longest tiles first (LPT), stable for reproducible schedules
    std::stable_sort(tl.begin(), tl.end(),
                     [](const std::pair<double, int4>& a, const std::pair<double, int4>& b) {
                         return a.first > b.first;
                     });
    std::vector<int4> tiles(tl.size());
    for (size_t i = 0; i < tl.size(); ++i) tiles[i] = tl[i].second;
    ssam_batch_plan_s::Group g;
    g.ndim = 2;
    g.storage = storage;
    g.ntiles = static_cast<int>(tiles.size());
    P->groups.push_back(g);  // owned by the plan from here on (freed on failure too)
    ssam_batch_plan_s::Group& gg = P->groups.back();
    if (hipMalloc(&gg.d_items, h.size() * sizeof(BatchItem2D)) != hipSuccess ||
        hipMalloc(&gg.d_tiles, std::max<size_t>(1, tiles.size()) * sizeof(int4)) != hipSuccess ||
        hipMalloc(&gg.d_taps, std::max<size_t>(1, taps.size()) * sizeof(A)) != hipSuccess)
        return SSAM_E_HIP;
    g = gg;
    if (hipMemcpy(g.d_items, h.data(), h.size() * sizeof(BatchItem2D), hipMemcpyHostToDevice) !=
            hipSuccess ||
        (tiles.size() && hipMemcpy(g.d_tiles, tiles.data(), tiles.size() * sizeof(int4),
                                   hipMemcpyHostToDevice) != hipSuccess) ||
        (taps.size() &&
         hipMemcpy(g.d_taps, taps.data(), taps.size() * sizeof(A), hipMemcpyHostToDevice) != hipSuccess))
        return SSAM_E_HIP;
    P->tiles2d += g.ntiles;
    return SSAM_OK;
}

template <typename S>
int build_group3d(ssam_batch_plan_s* P, const std::vector<int>& ids, const ssam_batch_item* items,
                  int storage) {
    constexpr int V = Lanes3Db<S>::V;
    std::vector<BatchItem3D> h(ids.size());
    std::vector<S> taps;
    std::vector<std::pair<double, int4>> tl;
    const int zseg = 16;
    int steps = 0;
    for (size_t k = 0; k < ids.size(); ++k) {
        const ssam_batch_item& it = items[ids[k]];
        BatchItem3D& b = h[k];
        b.src = it.src;
        b.dst = it.dst;
        b.work = it.work;
        b.kind = it.shape == SSAM_SHAPE_FACE_EDGE ? 1 : 0;
        b.steps = it.steps;
        if (steps == 0) steps = it.steps;
        if (it.steps != steps) return SSAM_E_UNSUPPORTED;  // one step count per group
        b.taps_off = static_cast<int>(taps.size());
        const int nt = b.kind ? 19 : 7;
        for (int t = 0; t < nt; ++t) taps.push_back(static_cast<S>(it.coeffs[t]));
        Params3D& p = b.p;
        p.nx = static_cast<int>(it.nx);
        p.ny = static_cast<int>(it.ny);
        p.nz = static_cast<int>(it.nz);
        p.sy = it.nx;
        p.sz = it.nx * it.ny;
        p.nstrips = static_cast<int>((it.nx + kWave * V - 1) / (kWave * V));
        p.nyb = static_cast<int>((it.ny + kNY3 - 1) / kNY3);
        p.zseg = static_cast<int>(std::min<int64_t>(zseg, it.nz));
        const int nzs = (p.nz + p.zseg - 1) / p.zseg;
        p.ntiles = p.nstrips * p.nyb * nzs;
        p.nblocks = 0;
        p.vec_ok = (it.nx % V == 0) && (reinterpret_cast<uintptr_t>(it.src) % 16 == 0) &&
                   (reinterpret_cast<uintptr_t>(it.dst) % 16 == 0) &&
                   (!it.work || reinterpret_cast<uintptr_t>(it.work) % 16 == 0);
        p.partials = nullptr;
        p.res_scale = 0.0;
        for (int z = 0; z < nzs; ++z)
            for (int y = 0; y < p.nyb; ++y)
                for (int s = 0; s < p.nstrips; ++s) {
                    const int w = std::min<int>(kWave * V, p.nx - s * kWave * V);
                    const int planes = std::min(p.zseg, p.nz - z * p.zseg);
                    tl.push_back({static_cast<double>(w) * (planes + 2) * nt,
                                  int4{static_cast<int>(k), s, y, z * p.zseg}});
                }
        P->point_updates += it.nx * it.ny * it.nz * it.steps;
    }
    std::stable_sort(tl.begin(), tl.end(),
                     [](const std::pair<double, int4>& a, const std::pair<double, int4>& b) {
                         return a.first > b.first;
                     });
    std::vector<int4> tiles(tl.size());
    for (size_t i = 0; i < tl.size(); ++i) tiles[i] = tl[i].second;
    ssam_batch_plan_s::Group g;
    g.ndim = 3;
    g.storage = storage;
    g.steps = steps;
    g.ntiles = static_cast<int>(tiles.size());
    P->groups.push_back(g);  // owned by the plan from here on (freed on failure too)
    ssam_batch_plan_s::Group& gg = P->groups.back();
    if (hipMalloc(&gg.d_items, h.size() * sizeof(BatchItem3D)) != hipSuccess ||
        hipMalloc(&gg.d_tiles, std::max<size_t>(1, tiles.size()) * sizeof(int4)) != hipSuccess ||
        hipMalloc(&gg.d_taps, std::max<size_t>(1, taps.size()) * sizeof(S)) != hipSuccess)
        return SSAM_E_HIP;
    g = gg;
    if (hipMemcpy(g.d_items, h.data(), h.size() * sizeof(BatchItem3D), hipMemcpyHostToDevice) !=
            hipSuccess ||
        (tiles.size() && hipMemcpy(g.d_tiles, tiles.data(), tiles.size() * sizeof(int4),
                                   hipMemcpyHostToDevice) != hipSuccess) ||
        (taps.size() &&
         hipMemcpy(g.d_taps, taps.data(), taps.size() * sizeof(S), hipMemcpyHostToDevice) != hipSuccess))
        return SSAM_E_HIP;
    P->tiles3d += static_cast<int64_t>(g.ntiles) * steps;
    return SSAM_OK;
}

void free_plan(ssam_batch_plan_s* P) {
    if (!P) return;
    if (P->exec) (void)hipGraphExecDestroy(P->exec);
    if (P->graph) (void)hipGraphDestroy(P->graph);
    if (P->cap) (void)hipStreamDestroy(P->cap);
    for (auto& g : P->groups) {
        if (g.d_items) (void)hipFree(g.d_items);
        if (g.d_tiles) (void)hipFree(g.d_tiles);
        if (g.d_taps) (void)hipFree(g.d_taps);
    }
    if (P->d_counters) (void)hipFree(P->d_counters);
    delete P;
}

// Every launch of one batch run, in order, on stream st.
int enqueue(ssam_batch_plan_s* P, hipStream_t st) {
    hipLaunchKernelGGL(reset_counters_kernel, dim3((P->ncounters + 255) / 256), dim3(256), 0, st,
                       static_cast<int*>(P->d_counters), P->ncounters);
    if (hipGetLastError() != hipSuccess) return SSAM_E_HIP;
    int ctr = 0;
    int* counters = static_cast<int*>(P->d_counters);
    for (auto& g : P->groups) {
        if (g.ndim != 2) continue;
        const int blocks = std::max(1, std::min(P->persistent_blocks,
                                                (g.ntiles + kBatchWavesPerBlock - 1) / kBatchWavesPerBlock));
        const dim3 grid(blocks), block(kWave * kBatchWavesPerBlock);
        const auto* items = static_cast<const BatchItem2D*>(g.d_items);
        const auto* tiles = static_cast<const int4*>(g.d_tiles);
        switch (g.storage) {
            case SSAM_F32:
                hipLaunchKernelGGL(batch2d_kernel<float>, grid, block, 0, st, items, tiles, g.ntiles,
                                   counters + ctr, static_cast<const float*>(g.d_taps));
                break;
            case SSAM_F64:
                hipLaunchKernelGGL(batch2d_kernel<double>, grid, block, 0, st, items, tiles, g.ntiles,
                                   counters + ctr, static_cast<const double*>(g.d_taps));
                break;
            case SSAM_BF16:
                hipLaunchKernelGGL(batch2d_kernel<bf16_t>, grid, block, 0, st, items, tiles, g.ntiles,
                                   counters + ctr, static_cast<const float*>(g.d_taps));
                break;
            default:
                hipLaunchKernelGGL(batch2d_kernel<f16_t>, grid, block, 0, st, items, tiles, g.ntiles,
                                   counters + ctr, static_cast<const float*>(g.d_taps));
                break;
        }
        if (hipGetLastError() != hipSuccess) return SSAM_E_HIP;
        ++ctr;
    }
    for (size_t i = 0; i < P->fallback.size(); ++i) {
        const ssam_batch_item& it = P->fallback[i];
        int rc;
        if (it.ndim == 2) {
            ssam_stencil2d_desc d{};
            d.storage = it.storage;
            d.shape = it.shape;
            d.fw = d.fh = 2 * it.radius + 1;
            d.steps = it.steps;
            d.boundary = SSAM_BOUNDARY_ZERO;
            d.coeffs = P->fallback_taps[i].data();
            rc = ssam_stencil2d(&d, it.src, it.dst, it.nx, it.ny, it.pitch, it.work, st);
        } else {
            ssam_stencil3d_desc d{};
            d.storage = it.storage;
            d.shape = it.shape;
            d.radius = it.radius;
            d.steps = it.steps;
            d.boundary = SSAM_BOUNDARY_DIRICHLET0;
            d.coeffs = P->fallback_taps[i].data();
            rc = ssam_stencil3d(&d, it.src, it.dst, it.nx, it.ny, it.nz, it.work, st);
        }
        if (rc != SSAM_OK) return rc;
    }
    int maxsteps = 0;
    for (auto& g : P->groups)
        if (g.ndim == 3) maxsteps = std::max(maxsteps, g.steps);
    for (int s = 0; s < maxsteps; ++s) {
        int c3 = ctr;
        for (auto& g : P->groups) {
            if (g.ndim != 3) continue;
            if (s < g.steps) {
                const int blocks = std::max(1, std::min(P->persistent_blocks,
                                                        (g.ntiles + kBatchWavesPerBlock - 1) / kBatchWavesPerBlock));
                const dim3 grid(blocks), block(kWave * kBatchWavesPerBlock);
                const auto* items = static_cast<const BatchItem3D*>(g.d_items);
                const auto* tiles = static_cast<const int4*>(g.d_tiles);
                if (g.storage == SSAM_F64)
                    hipLaunchKernelGGL(batch3d_kernel<double>, grid, block, 0, st, items, tiles,
                                       g.ntiles, counters + c3, static_cast<const double*>(g.d_taps), s);
                else
                    hipLaunchKernelGGL(batch3d_kernel<float>, grid, block, 0, st, items, tiles,
                                       g.ntiles, counters + c3, static_cast<const float*>(g.d_taps), s);
                if (hipGetLastError() != hipSuccess) return SSAM_E_HIP;
            }
            ++c3;
        }
        ctr = c3;
    }
    return SSAM_OK;
}

}  // namespace

extern "C" {

int ssam_batch_plan_create(const ssam_batch_item* items, int n, ssam_batch_plan* out) {
    if (!items || n < 1 || !out) return SSAM_E_INVALID;
    *out = nullptr;
    // validate and group
    std::vector<int> g2[4], g3[2];
    std::vector<int> fb;
    for (int i = 0; i < n; ++i) {
        const ssam_batch_item& it = items[i];
        if (!it.src || !it.dst || !it.coeffs || it.steps < 1 || it.radius < 1) return SSAM_E_INVALID;
        const int es = elem_size_b(it.storage);
        if (!es) return SSAM_E_INVALID;
        if (it.ndim == 2) {
            if (it.nx < 1 || it.ny < 1 || it.pitch < it.nx || it.nz != 1) return SSAM_E_INVALID;
            if (it.nx > (1 << 30) || it.ny > (1 << 30)) return SSAM_E_UNSUPPORTED;
            if (it.shape != SSAM_SHAPE_STAR && it.shape != SSAM_SHAPE_BOX) return SSAM_E_INVALID;
            const int64_t bytes = ((it.ny - 1) * it.pitch + it.nx) * es;
            const char* a = static_cast<const char*>(it.src);
            const char* b = static_cast<const char*>(it.dst);
            if (a < b + bytes && b < a + bytes) return SSAM_E_INVALID;
            const int k = kind2d(it.shape == SSAM_SHAPE_STAR ? kStar : kBox, it.radius);
            if (k == k2None || it.steps != 1) fb.push_back(i);
            else g2[it.storage].push_back(i);
        } else if (it.ndim == 3) {
            if (it.nx < 1 || it.ny < 1 || it.nz < 1) return SSAM_E_INVALID;
            if (it.storage != SSAM_F32 && it.storage != SSAM_F64) return SSAM_E_UNSUPPORTED;
            if (it.steps > 1 && !it.work) return SSAM_E_INVALID;
            const int64_t bytes = it.nx * it.ny * it.nz * es;
            const char* a = static_cast<const char*>(it.src);
            const char* b = static_cast<const char*>(it.dst);
            if (a < b + bytes && b < a + bytes) return SSAM_E_INVALID;
            if (it.radius == 1 && (it.shape == SSAM_SHAPE_STAR || it.shape == SSAM_SHAPE_FACE_EDGE))
                g3[it.storage == SSAM_F64 ? 1 : 0].push_back(i);
            else
                fb.push_back(i);
        } else {
            return SSAM_E_INVALID;
        }
    }
    auto* P = new (std::nothrow) ssam_batch_plan_s();
    if (!P) return SSAM_E_NOMEM;
    (void)hipGetDevice(&P->device);
    int rc = SSAM_OK;
    if (!g2[SSAM_F32].empty()) rc = build_group2d<float>(P, g2[SSAM_F32], items, SSAM_F32);
    if (rc == SSAM_OK && !g2[SSAM_F64].empty()) rc = build_group2d<double>(P, g2[SSAM_F64], items, SSAM_F64);
    if (rc == SSAM_OK && !g2[SSAM_BF16].empty()) rc = build_group2d<bf16_t>(P, g2[SSAM_BF16], items, SSAM_BF16);
    if (rc == SSAM_OK && !g2[SSAM_F16].empty()) rc = build_group2d<f16_t>(P, g2[SSAM_F16], items, SSAM_F16);
    // 3D groups: split by step count inside build_group3d (one count per group)
    for (int s = 0; s < 2 && rc == SSAM_OK; ++s) {
        if (g3[s].empty()) continue;
        std::vector<std::vector<int>> bysteps;
        std::vector<int> keys;
        for (int i : g3[s]) {
            auto pos = std::find(keys.begin(), keys.end(), items[i].steps);
            if (pos == keys.end()) {
                keys.push_back(items[i].steps);
                bysteps.push_back({i});
            } else {
                bysteps[pos - keys.begin()].push_back(i);
            }
        }
        for (auto& ids : bysteps) {
            rc = s ? build_group3d<double>(P, ids, items, SSAM_F64) : build_group3d<float>(P, ids, items, SSAM_F32);
            if (rc != SSAM_OK) break;
        }
    }
    for (int i : fb) {
        const ssam_batch_item& it = items[i];
        P->fallback.push_back(it);
        const int nt = it.ndim == 2 ? (2 * it.radius + 1) * (2 * it.radius + 1)
                                    : (it.shape == SSAM_SHAPE_FACE_EDGE ? 19 : 6 * it.radius + 1);
        P->fallback_taps.emplace_back(it.coeffs, it.coeffs + nt);
        P->point_updates += it.nx * it.ny * it.nz * it.steps;
    }
    if (rc == SSAM_OK) {
        // one counter per 2D launch and per (3D group, step) slot of enqueue()
        int n2 = 0, n3 = 0, maxsteps = 0;
        for (auto& g : P->groups) {
            if (g.ndim == 2) ++n2;
            else {
                ++n3;
                maxsteps = std::max(maxsteps, g.steps);
            }
        }
        P->ncounters = std::max(1, n2 + n3 * maxsteps);
        if (hipMalloc(&P->d_counters, P->ncounters * sizeof(int)) != hipSuccess) rc = SSAM_E_HIP;
    }
    if (rc != SSAM_OK) {
        free_plan(P);
        return rc;
    }
    *out = P;
    return SSAM_OK;
}

int ssam_batch_run(ssam_batch_plan plan, void* stream, uint32_t flags) {
    if (!plan) return SSAM_E_INVALID;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (flags & SSAM_BATCH_EAGER) return enqueue(plan, st);
    if (!plan->exec) {
        // capture the whole run once on a private stream (default priority)
        if (!plan->cap && hipStreamCreateWithFlags(&plan->cap, hipStreamNonBlocking) != hipSuccess)
            return SSAM_E_HIP;
        if (hipStreamBeginCapture(plan->cap, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return SSAM_E_HIP;
        const int rc = enqueue(plan, plan->cap);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(plan->cap, &g);
        if (rc != SSAM_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess || !g) return SSAM_E_HIP;
        plan->graph = g;
        if (hipGraphInstantiate(&plan->exec, g, nullptr, nullptr, 0) != hipSuccess) {
            plan->exec = nullptr;
            return SSAM_E_HIP;
        }
    }
    return hipGraphLaunch(plan->exec, st) == hipSuccess ? SSAM_OK : SSAM_E_HIP;
}

int ssam_batch_plan_info(ssam_batch_plan plan, int64_t* tiles_2d, int64_t* tiles_3d,
                         int64_t* point_updates) {
    if (!plan) return SSAM_E_INVALID;
    if (tiles_2d) *tiles_2d = plan->tiles2d;
    if (tiles_3d) *tiles_3d = plan->tiles3d;
    if (point_updates) *point_updates = plan->point_updates;
    return SSAM_OK;
}

void ssam_batch_plan_destroy(ssam_batch_plan plan) { free_plan(plan); }

}  // extern "C"
