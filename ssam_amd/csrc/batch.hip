// batch.hip -- many grids of uneven sizes and mixed stencil shapes: the plan,
// its tile tables, the launch sequence captured in a hipGraph, and the tuner
// (the kernels are in batch2d.hip and batch3d.hip).
//
// No counterpart in the reference (one grid, one launch per harness call).
//
//  * Plan: every 2D item is cut into tiles (column strip x segment of rows) of the
//    same geometry the single-grid launch uses; 3D items into tiles of (x-strip x
//    y-block x z-segment), one or two sweeps per pass.
//  * Launch forms: an ordinary grid with one tile per wave over a table sorted by
//    decreasing tile cost (the hardware workgroup dispatcher is the queue; the
//    default), or a persistent grid of as many waves as the device holds, which
//    drains 8 work queues (one per XCD, each item's tiles in memory order, cut
//    into runs packed to equal cost): a wave reads its XCD id, takes the next tile
//    of that queue (one atomic from lane 0), steals from the other queues when its
//    own is empty, and leaves when all are empty; the result never depends on which
//    XCD runs what.  2D items run one launch per stencil kind or one for all kinds.
//  * Graph: counter reset (a fill kernel, no memset), the 2D launches, the 3D
//    passes and any per-item fallback launches are captured once on a private
//    stream and replayed on the caller's stream.
//  * Tuner: ssam_batch_plan_tune times launch forms, tile heights per 2D shape,
//    3D tile shapes (temporal degree included), occupancy and queue affinity on
//    the plan's own buffers and keeps the fastest.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "batch_common.hpp"

namespace ssam {

__global__ void reset_counters_kernel(int* __restrict__ c, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c[i] = 0;
}

}  // namespace ssam

using namespace ssam;

// ------------------------------------------------------------------ plan
struct ssam_batch_plan_s {
    int device = 0;
    struct Group {  // 2D: one launch per kind or for all kinds; 3D: one launch per pass
        int ndim = 2;
        int storage = 0;
        std::vector<int> ids;  // indices into items
        void* d_items = nullptr;
        void* d_tiles = nullptr;
        void* d_taps = nullptr;
        int ntiles = 0;
        QueueSet qs{};    // 3D
        struct Sub {      // 2D: one persistent launch per kind (split) or one for all (fused)
            int kind = -1;
            int tile_off = 0, ntiles = 0;
            QueueSet qs{};
            int blocks = 1;
        };
        std::vector<Sub> subs;
        int steps = 1;    // 3D
        int kind = 0;     // 3D stencil: 0 7-point star, 1 19-point face+edge
        int blocks = 1;   // persistent grid: resident workgroups of its kernel
    };
    std::vector<ssam_batch_item> items;       // copies; coeffs point into taps_host
    std::vector<std::vector<double>> taps_host;
    std::vector<Group> groups;
    std::vector<int> fallback;                // indices into items
    int seg2d[kNumKind2D] = {kDefaultSeg2D, kDefaultSeg2D, kDefaultSeg2D, kDefaultSeg2D};
    int var3d = kDefaultVar3D, zseg3d = kDefaultZseg3D;
    int aff2d = 1, aff3d = 0;  // queue affinity (QueueSet::affinity)
    int alt3d = 1;             // 3D sweeps alternate direction (QueueSet::alternate)
    int split2d = 1;           // 2D: one launch per kind (1) or one for all kinds (0)
    int mfma2d = 3;            // BF16/FP16 boxes on the matrix cores: bit 0 5x5, bit 1 11x11
    int wps2d = 0, wps3d = 1;  // persistent grid: waves per SIMD (0: as many as fit)
    int grid2d = 1, grid3d = 1;  // one tile per wave of an ordinary grid (1) or persistent (0)
    int ncounters = 0;
    void* d_counters = nullptr;
    int64_t tiles2d = 0, tiles3d = 0, point_updates = 0;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

using Plan = ssam_batch_plan_s;

bool two_byte(int storage) { return storage == SSAM_BF16 || storage == SSAM_F16; }

// Tile shape of a 3D group: the plan's, except that BF16 / F16 groups run a
// multi-sweep shape (the only ones instantiated for 2-byte storage and for the
// FP32 state between their passes: three sweeps when the plan's is, else two),
// and 19-point groups take two sweeps where the plan's shape has three.
int var3d_of(const Plan* P, const Plan::Group& g) {
    if (P->var3d == kVarT3) return g.kind ? kVarT2 : kVarT3;
    return two_byte(g.storage) ? kVarT2 : P->var3d;
}

// 3D launch form in effect: the multi-sweep shapes always run as ordinary grids.
bool grid3d_on(const Plan* P, const Plan::Group& g) { return P->grid3d || kVar3D[var3d_of(P, g)].t >= 2; }

int elem_size_b(int dtype) {
    switch (dtype) {
        case SSAM_F32: return 4;
        case SSAM_F64: return 8;
        case SSAM_BF16:
        case SSAM_F16: return 2;
        default: return 0;
    }
}

int env_int(const char* name, int dflt) {
    const char* v = std::getenv(name);
    return v && *v ? std::atoi(v) : dflt;
}

// The persistent grid of `fn`: the workgroups the device holds at once, at
// most `wps` per CU (one wave per SIMD each; 0: no cap).
int resident_blocks(const void* fn, int wps) {
    const int forced = env_int("SSAM_BATCH_BLOCKS", 0);
    if (forced > 0) return forced;
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, kWave * kBatchWavesPerBlock, 0) !=
            hipSuccess ||
        per < 1)
        per = 1;
    if (wps > 0 && wps < per) per = wps;
    return cus * per;
}

// One item's tiles in memory order with a per-tile cost estimate.
struct ItemTiles {
    std::vector<int4> tiles;
    std::vector<double> cost;
};

// Pack items' tiles into kQueues queues: each item is cut into runs of
// consecutive tiles (at most 1/32 of the total cost), runs are dealt to the least
// loaded queue largest first, and inside a queue items go in decreasing per-tile
// cost, each item's runs in memory order.
void pack_queues(const std::vector<ItemTiles>& per_item, std::vector<int4>* tiles, QueueSet* qs) {
    struct Run { int item; size_t b, e; double cost; };
    double total = 0;
    for (auto& it : per_item)
        for (double c : it.cost) total += c;
    const double cap = total / (kQueues * 4) > 0 ? total / (kQueues * 4) : 1.0;
    std::vector<Run> runs;
    for (int i = 0; i < static_cast<int>(per_item.size()); ++i) {
        const ItemTiles& it = per_item[i];
        size_t b = 0;
        double acc = 0;
        for (size_t k = 0; k < it.tiles.size(); ++k) {
            acc += it.cost[k];
            if (acc >= cap || k + 1 == it.tiles.size()) {
                runs.push_back({i, b, k + 1, acc});
                b = k + 1;
                acc = 0;
            }
        }
    }
    std::vector<int> order(runs.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = static_cast<int>(i);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return runs[a].cost > runs[b].cost; });
    std::vector<std::vector<int>> bins(kQueues);
    double load[kQueues] = {0};
    for (int r : order) {
        int q = 0;
        for (int k = 1; k < kQueues; ++k)
            if (load[k] < load[q]) q = k;
        bins[q].push_back(r);
        load[q] += runs[r].cost;
    }
    auto tile_cost = [&](int r) { return per_item[runs[r].item].cost[runs[r].b]; };
    tiles->clear();
    for (int q = 0; q < kQueues; ++q) {
        std::vector<int>& bq = bins[q];
        std::sort(bq.begin(), bq.end(), [&](int a, int b) {
            const double ca = tile_cost(a), cb = tile_cost(b);
            if (ca != cb) return ca > cb;
            return runs[a].item != runs[b].item ? runs[a].item < runs[b].item : runs[a].b < runs[b].b;
        });
        qs->beg[q] = static_cast<int>(tiles->size());
        for (int r : bq)
            for (size_t k = runs[r].b; k < runs[r].e; ++k) tiles->push_back(per_item[runs[r].item].tiles[k]);
    }
    qs->beg[kQueues] = static_cast<int>(tiles->size());
}

// One list for a grid-mode launch: all tiles by decreasing cost (the hardware
// dispatches workgroups in order, so the longest tiles start first), each item's
// tiles of equal cost in memory order.  Every queue bound points at the end.
void pack_single(const std::vector<ItemTiles>& per_item, std::vector<int4>* tiles, QueueSet* qs) {
    struct Ref { double cost; int item; int k; };
    std::vector<Ref> all;
    for (int i = 0; i < static_cast<int>(per_item.size()); ++i)
        for (int k = 0; k < static_cast<int>(per_item[i].tiles.size()); ++k)
            all.push_back({per_item[i].cost[k], i, k});
    std::stable_sort(all.begin(), all.end(), [](const Ref& a, const Ref& b) { return a.cost > b.cost; });
    tiles->clear();
    for (const Ref& r : all) tiles->push_back(per_item[r.item].tiles[r.k]);
    qs->beg[0] = 0;
    for (int q = 1; q <= kQueues; ++q) qs->beg[q] = static_cast<int>(tiles->size());
}

template <typename T>
int upload(void** dptr, const std::vector<T>& h) {
    if (hipMalloc(dptr, std::max<size_t>(1, h.size()) * sizeof(T)) != hipSuccess) {
        *dptr = nullptr;
        return SSAM_E_HIP;
    }
    if (!h.empty() && hipMemcpy(*dptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        return SSAM_E_HIP;
    return SSAM_OK;
}

template <typename S, int R, int SHAPE>
void tiles_for_2d(const ssam_batch_item& it, int idx, int seg_rows, Params2D* p, ItemTiles* out) {
    using I = Inst2D<S, R, SHAPE>;
    p->width = static_cast<int>(it.nx);
    p->height = static_cast<int>(it.ny);
    p->pitch = it.pitch;
    p->pitch_out = it.pitch;
    p->nstrips = static_cast<int>((it.nx + I::G::OW - 1) / I::G::OW);
    p->seg_rows = seg_rows;
    p->boundary = kZero;
    p->taps = nullptr;  // the item's taps reach the tile as a pointer into the table
    p->vec_ok = (reinterpret_cast<uintptr_t>(it.src) % sizeof(S) == 0) &&
                (reinterpret_cast<uintptr_t>(it.dst) % sizeof(S) == 0);
    const int nsegs = (p->height + seg_rows - 1) / seg_rows;
    p->ntiles = p->nstrips * nsegs;
    p->nblocks = 0;
    const double taps = SHAPE == kStar ? 4 * R + 1 : (2 * R + 1) * (2 * R + 1);
    for (int sg = 0; sg < nsegs; ++sg) {
        const int rows = std::min(seg_rows, p->height - sg * seg_rows);
        for (int st = 0; st < p->nstrips; ++st) {
            out->tiles.push_back(int4{idx, st, sg, 0});
            out->cost.push_back(static_cast<double>(rows + 2 * R) * (taps + 8.0));
        }
    }
}

// Tiles of the matrix-core boxes: 32-column strips, rows in whole ring periods
// (64).
void tiles_for_mfma(const ssam_batch_item& it, int idx, int R, int seg_rows, ParamsM* p, ItemTiles* out) {
    constexpr int OW = 32;
    p->width = static_cast<int>(it.nx);
    p->height = static_cast<int>(it.ny);
    p->pitch = it.pitch;
    p->nstrips = static_cast<int>((it.nx + OW - 1) / OW);
    p->seg_rows = std::max(64, (seg_rows + 63) / 64 * 64);
    const int nsegs = (p->height + p->seg_rows - 1) / p->seg_rows;
    p->ntiles = p->nstrips * nsegs;
    p->nblocks = 0;
    for (int sg = 0; sg < nsegs; ++sg) {
        const int rows = std::min(p->seg_rows, p->height - sg * p->seg_rows);
        for (int st = 0; st < p->nstrips; ++st) {
            out->tiles.push_back(int4{idx, st, sg, 0});
            // a quarter of the VALU kernel's cost per point (measured ratio on
            // the 11x11 box)
            const double taps = (2.0 * R + 1) * (2.0 * R + 1);
            out->cost.push_back(static_cast<double>(rows + 2 * R) * (taps + 8.0) * 0.25 * OW / 248.0);
        }
    }
}

void free_group(Plan::Group& g) {
    if (g.d_items) (void)hipFree(g.d_items);
    if (g.d_tiles) (void)hipFree(g.d_tiles);
    if (g.d_taps) (void)hipFree(g.d_taps);
    g.d_items = g.d_tiles = g.d_taps = nullptr;
}

// (Re)build the device tables of a 2D group with the plan's tile heights: one
// queue set over all tiles (fused) or one per launch kind (split), the kinds'
// tiles one after the other in one table.
template <typename S>
int build_group2d(Plan* P, Plan::Group& g) {
    using A = typename Storage<S>::acc;
    std::vector<BatchItem2D> h(g.ids.size());
    std::vector<A> taps;
    std::vector<ItemTiles> per(g.ids.size());
    for (size_t k = 0; k < g.ids.size(); ++k) {
        const ssam_batch_item& it = P->items[g.ids[k]];
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
        const int seg = P->seg2d[b.kind];
        // 5x5 and 11x11 boxes in BF16/FP16 go to the matrix cores
        // (matrix-core edge tiles address the grid through one raw buffer: < 2 GiB)
        if (((b.kind == k2Box2 && (P->mfma2d & 1)) || (b.kind == k2Box5 && (P->mfma2d & 2))) && kHalf2D<S> &&
            ((reinterpret_cast<uintptr_t>(it.src) | reinterpret_cast<uintptr_t>(it.dst)) % 2) == 0 &&
            ((it.ny - 1) * it.pitch + it.nx) * 2 < (1LL << 31)) {
            b.kind = b.kind == k2Box2 ? k2Box2M : k2Box5M;
            tiles_for_mfma(it, idx, R, seg, &b.pm, &per[k]);
            continue;
        }
        switch (b.kind) {
            case k2Star1: tiles_for_2d<S, 1, kStar>(it, idx, seg, &b.p, &per[k]); break;
            case k2Box1: tiles_for_2d<S, 1, kBox>(it, idx, seg, &b.p, &per[k]); break;
            case k2Box2: tiles_for_2d<S, 2, kBox>(it, idx, seg, &b.p, &per[k]); break;
            default: tiles_for_2d<S, 5, kBox>(it, idx, seg, &b.p, &per[k]); break;
        }
    }
    std::vector<int4> tiles;
    g.subs.clear();
    const int nkinds = P->split2d ? kNumLaunch2D : 1;
    for (int kd = 0; kd < nkinds; ++kd) {
        std::vector<ItemTiles> part(per.size());
        bool any = false;
        for (size_t k = 0; k < per.size(); ++k)
            if (!P->split2d || h[k].kind == kd) {
                part[k] = per[k];
                any = any || !per[k].tiles.empty();
            }
        if (!any) continue;
        Plan::Group::Sub sb;
        sb.kind = P->split2d ? kd : -1;
        std::vector<int4> t;
        if (P->grid2d) pack_single(part, &t, &sb.qs);
        else pack_queues(part, &t, &sb.qs);
        sb.qs.affinity = P->aff2d;
        sb.qs.alternate = 0;
        sb.tile_off = static_cast<int>(tiles.size());
        sb.ntiles = static_cast<int>(t.size());
        sb.blocks = P->grid2d ? (sb.ntiles + kBatchWavesPerBlock - 1) / kBatchWavesPerBlock
                              : resident_blocks(batch2d_kernel_fn(g.storage, sb.kind, false), P->wps2d);
        tiles.insert(tiles.end(), t.begin(), t.end());
        g.subs.push_back(sb);
    }
    free_group(g);
    g.ntiles = static_cast<int>(tiles.size());
    int rc = upload(&g.d_items, h);
    if (rc == SSAM_OK) rc = upload(&g.d_tiles, tiles);
    if (rc == SSAM_OK) rc = upload(&g.d_taps, taps);
    return rc;
}

// (Re)build the device tables of a 3D group with the plan's tile shape.
template <typename S>
int build_group3d(Plan* P, Plan::Group& g) {
    using A = typename Storage<S>::acc;
    const int variant = var3d_of(P, g), zseg = P->zseg3d;
    std::vector<BatchItem3D> h(g.ids.size());
    std::vector<A> taps;
    std::vector<ItemTiles> per(g.ids.size());
    for (size_t k = 0; k < g.ids.size(); ++k) {
        const ssam_batch_item& it = P->items[g.ids[k]];
        BatchItem3D& b = h[k];
        b.src = it.src;
        b.dst = it.dst;
        b.work = it.work;
        b.taps_off = static_cast<int>(taps.size());
        b.var = variant == kVarPair3D ? pair_choice(it.nx, sizeof(A)) : variant;
        b.steps = passes3d(static_cast<int>(it.steps), variant);
        b.fp32_state = two_byte(g.storage) && b.steps > 1;
        const int T = kVar3D[variant].t;
        const int V = var3d_v(b.var, sizeof(A)), NY = kVar3D[b.var].ny;
        const int OW = (kWave - 2 * var3d_hl(variant, V)) * V;  // halo lanes: a column per sweep
        const int nt = g.kind ? 19 : 7;
        for (int t = 0; t < nt; ++t) taps.push_back(static_cast<A>(it.coeffs[t]));
        Grid3D& gr = b.g;
        gr.nx = static_cast<int>(it.nx);
        gr.ny = static_cast<int>(it.ny);
        gr.nz = static_cast<int>(it.nz);
        gr.sy = it.nx;
        gr.sz = it.nx * it.ny;
        b.nstrips = static_cast<int>((it.nx + OW - 1) / OW);
        b.nyb = static_cast<int>((it.ny + NY - 1) / NY);
        const uintptr_t al = static_cast<uintptr_t>(V) * sizeof(S);
        const uintptr_t alw = static_cast<uintptr_t>(V) * (b.fp32_state ? sizeof(float) : sizeof(S));
        b.vec = (it.nx % V == 0) && (reinterpret_cast<uintptr_t>(it.src) % al == 0) &&
                (reinterpret_cast<uintptr_t>(it.dst) % al == 0) &&
                (!it.work || reinterpret_cast<uintptr_t>(it.work) % alw == 0);
        const int zs_len = std::min<int>(zseg, gr.nz);
        for (int z0 = 0; z0 < gr.nz; z0 += zs_len) {
            const int planes = std::min(zs_len, gr.nz - z0);
            for (int y = 0; y < b.nyb; ++y)
                for (int s = 0; s < b.nstrips; ++s) {
                    const int rows = std::min(NY, gr.ny - y * NY);
                    per[k].tiles.push_back(int4{static_cast<int>(k), s, y, z0 | (planes << 20)});
                    // rows loaded x planes streamed (+2 halo planes per sweep) x taps
                    // x sweeps
                    per[k].cost.push_back(static_cast<double>(rows + 2 * T) * (planes + 2 * T) * nt * T);
                }
        }
    }
    std::vector<int4> tiles;
    if (grid3d_on(P, g)) pack_single(per, &tiles, &g.qs);
    else pack_queues(per, &tiles, &g.qs);
    g.qs.affinity = P->aff3d;
    g.qs.alternate = P->alt3d;
    free_group(g);
    g.ntiles = static_cast<int>(tiles.size());
    g.blocks = grid3d_on(P, g)
                   ? (g.ntiles + kBatchWavesPerBlock - 1) / kBatchWavesPerBlock
                   : resident_blocks(batch3d_kernel_fn(g.storage, g.storage, variant, g.kind, false), P->wps3d);
    int rc = upload(&g.d_items, h);
    if (rc == SSAM_OK) rc = upload(&g.d_tiles, tiles);
    if (rc == SSAM_OK) rc = upload(&g.d_taps, taps);
    return rc;
}

int build_group(Plan* P, Plan::Group& g) {
    if (g.ndim == 2) {
        switch (g.storage) {
            case SSAM_F32: return build_group2d<float>(P, g);
            case SSAM_F64: return build_group2d<double>(P, g);
            case SSAM_BF16: return build_group2d<bf16_t>(P, g);
            default: return build_group2d<f16_t>(P, g);
        }
    }
    switch (g.storage) {
        case SSAM_F64: return build_group3d<double>(P, g);
        case SSAM_F32: return build_group3d<float>(P, g);
        case SSAM_BF16: return build_group3d<bf16_t>(P, g);
        default: return build_group3d<f16_t>(P, g);
    }
}

int build_all(Plan* P) {
    P->tiles2d = P->tiles3d = 0;
    for (auto& g : P->groups) {
        const int rc = build_group(P, g);
        if (rc != SSAM_OK) return rc;
        if (g.ndim == 2) P->tiles2d += g.ntiles;
        else P->tiles3d += static_cast<int64_t>(g.ntiles) * g.steps;
    }
    return SSAM_OK;
}

void drop_graph(Plan* P) {
    if (P->exec) (void)hipGraphExecDestroy(P->exec);
    if (P->graph) (void)hipGraphDestroy(P->graph);
    P->exec = nullptr;
    P->graph = nullptr;
}

void free_plan(Plan* P) {
    if (!P) return;
    drop_graph(P);
    if (P->cap) (void)hipStreamDestroy(P->cap);
    for (auto& g : P->groups) free_group(g);
    if (P->d_counters) (void)hipFree(P->d_counters);
    delete P;
}

constexpr unsigned kRun2D = 1, kRun3D = 2, kRunFallback = 4, kRunAll = 7;

// The launches of one batch run, in order, on stream st (parts: kRun* mask).
int enqueue(Plan* P, hipStream_t st, unsigned parts = kRunAll) {
    hipLaunchKernelGGL(reset_counters_kernel, dim3((P->ncounters + 255) / 256), dim3(256), 0, st,
                       static_cast<int*>(P->d_counters), P->ncounters);
    if (hipGetLastError() != hipSuccess) return SSAM_E_HIP;
    int* counters = static_cast<int*>(P->d_counters);
    int ctr = 0;  // counter sets (kCounterSet ints each): 2D groups, then 3D (step, group)
    for (auto& g : P->groups) {
        if (g.ndim != 2) continue;
        // kNumLaunch2D counter sets per 2D group, whatever its form
        int* cnt = counters + kCounterSet * ctr;
        ctr += kNumLaunch2D;
        if (!(parts & kRun2D)) continue;
        for (size_t k = 0; k < g.subs.size(); ++k) {
            const Plan::Group::Sub& sb = g.subs[k];
            const int blocks = std::max(1, std::min(sb.blocks, (sb.ntiles + kBatchWavesPerBlock - 1) /
                                                                   kBatchWavesPerBlock));
            const void* items = g.d_items;
            const void* tiles = static_cast<const int4*>(g.d_tiles) + sb.tile_off;
            QueueSet qs = sb.qs;
            int* c = cnt + kCounterSet * k;
            const void* taps = g.d_taps;
            void* args[] = {&items, &tiles, &qs, &c, &taps};
            const void* fn = batch2d_kernel_fn(g.storage, sb.kind, P->grid2d != 0);
            if (hipLaunchKernel(fn, dim3(blocks), dim3(kWave * kBatchWavesPerBlock), args, 0, st) !=
                    hipSuccess ||
                hipGetLastError() != hipSuccess)
                return SSAM_E_HIP;
        }
    }
    if (parts & kRunFallback) {
        for (int i : P->fallback) {
            const ssam_batch_item& it = P->items[i];
            int rc;
            if (it.ndim == 2) {
                ssam_stencil2d_desc d{};
                d.storage = it.storage;
                d.shape = it.shape;
                d.fw = d.fh = 2 * it.radius + 1;
                d.steps = it.steps;
                d.boundary = SSAM_BOUNDARY_ZERO;
                d.coeffs = it.coeffs;
                rc = ssam_stencil2d(&d, it.src, it.dst, it.nx, it.ny, it.pitch, it.work, st);
            } else {
                ssam_stencil3d_desc d{};
                d.storage = it.storage;
                d.shape = it.shape;
                d.radius = it.radius;
                d.steps = it.steps;
                d.boundary = SSAM_BOUNDARY_DIRICHLET0;
                if (two_byte(it.storage)) d.flags = SSAM_FLAG_WORK_FP32;  // work: two FP32 grids
                d.coeffs = it.coeffs;
                rc = ssam_stencil3d(&d, it.src, it.dst, it.nx, it.ny, it.nz, it.work, st);
            }
            if (rc != SSAM_OK) return rc;
        }
    }
    // passes: one sweep each, two (the last pass of an odd sweep count runs one)
    // or three (four sweeps as 2+2, the rest twos and ones)
    int maxsteps = 0;
    for (auto& g : P->groups)
        if (g.ndim == 3) maxsteps = std::max(maxsteps, g.steps);
    for (int s = 0; s < maxsteps; ++s) {
        for (auto& g : P->groups) {
            if (g.ndim != 3) continue;
            int* cnt = counters + kCounterSet * ctr++;
            const int var = var3d_of(P, g);
            const int t = kVar3D[var].t;
            const int npass = passes3d(g.steps, var);
            if (!(parts & kRun3D) || s >= npass) continue;
            // the two-sweep kernel takes a flag for its one-sweep last pass; a
            // three-sweep shape launches the kernel of the pass's sweep count
            int single = t == 2 && (g.steps % 2 == 1) && s == npass - 1;
            int kvar = var;
            if (t == 3) {
                const int nsw = pass_sweeps3d(g.steps, var, s);
                kvar = nsw == 3 ? kVarT3 : (nsw == 2 ? kVarT3S2 : kVarT3S1);
            }
            // storage read and written by this pass (FP32 between the passes of a
            // BF16 / F16 group)
            int in_st = g.storage, out_st = g.storage;
            if (two_byte(g.storage) && npass > 1) {
                in_st = s == 0 ? g.storage : SSAM_F32;
                out_st = s == npass - 1 ? g.storage : SSAM_F32;
            }
            const int blocks = std::max(1, std::min(g.blocks, (g.ntiles + kBatchWavesPerBlock - 1) /
                                                                  kBatchWavesPerBlock));
            const BatchItem3D* items = static_cast<const BatchItem3D*>(g.d_items);
            const int4* tiles = static_cast<const int4*>(g.d_tiles);
            QueueSet qs = g.qs;
            int step = s;
            const void* taps = g.d_taps;
            void* args[] = {&items, &tiles, &qs, &cnt, &taps, &step, &single};
            const void* fn = batch3d_kernel_fn(in_st, out_st, kvar, g.kind, grid3d_on(P, g));
            if (!fn) return SSAM_E_UNSUPPORTED;
            if (hipLaunchKernel(fn, dim3(blocks), dim3(kWave * kBatchWavesPerBlock), args, 0, st) !=
                    hipSuccess ||
                hipGetLastError() != hipSuccess)
                return SSAM_E_HIP;
        }
    }
    return SSAM_OK;
}

// Mean time of `reps` eager runs of the given parts (after one warm-up run).
int time_parts(Plan* P, hipStream_t st, unsigned parts, int reps, float* ms) {
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
        if (a) (void)hipEventDestroy(a);
        return SSAM_E_HIP;
    }
    int rc = enqueue(P, st, parts);
    if (rc == SSAM_OK && hipEventRecord(a, st) != hipSuccess) rc = SSAM_E_HIP;
    for (int r = 0; r < reps && rc == SSAM_OK; ++r) rc = enqueue(P, st, parts);
    if (rc == SSAM_OK && (hipEventRecord(b, st) != hipSuccess || hipEventSynchronize(b) != hipSuccess))
        rc = SSAM_E_HIP;
    float t = 0.f;
    if (rc == SSAM_OK && hipEventElapsedTime(&t, a, b) != hipSuccess) rc = SSAM_E_HIP;
    *ms = t / reps;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return rc;
}

}  // namespace

extern "C" {

int ssam_batch_plan_create(const ssam_batch_item* items, int n, ssam_batch_plan* out) {
    if (!items || n < 1 || !out) return SSAM_E_INVALID;
    *out = nullptr;
    // validate and group
    std::vector<int> g2[4];
    std::vector<std::vector<int>> g3;  // one group per (storage, steps, shape)
    std::vector<int64_t> g3key;
    std::vector<int> fb;
    for (int i = 0; i < n; ++i) {
        const ssam_batch_item& it = items[i];
        if (!it.src || !it.dst || !it.coeffs || it.steps < 1 || it.radius < 1) return SSAM_E_INVALID;
        const int es = elem_size_b(it.storage);
        if (!es) return SSAM_E_INVALID;
        if (it.ndim == 2) {
            if (it.nx < 1 || it.ny < 1 || it.pitch < it.nx || it.nz != 1) return SSAM_E_INVALID;
            if (it.nx > (1 << 30) || it.ny > (1 << 30) || it.nx * es >= (1LL << 31)) return SSAM_E_UNSUPPORTED;
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
            if (it.steps > 1 && !it.work) return SSAM_E_INVALID;
            const int64_t bytes = it.nx * it.ny * it.nz * es;
            const char* a = static_cast<const char*>(it.src);
            const char* b = static_cast<const char*>(it.dst);
            if (a < b + bytes && b < a + bytes) return SSAM_E_INVALID;
            // (BF16 / F16: the FP32 state between passes must fit the buffer too)
            if (it.radius == 1 && (it.shape == SSAM_SHAPE_STAR || it.shape == SSAM_SHAPE_FACE_EDGE) &&
                grid3d_fits_buffer(it.nx, it.ny, it.nz, std::max(es, 4))) {
                // tile fields: first plane in 20 bits, plane count in 11
                if (it.nz >= (1 << 20) || it.nx > (1 << 30) || it.ny > (1 << 30))
                    return SSAM_E_UNSUPPORTED;
                const int64_t key = (static_cast<int64_t>(it.steps) * 2 +
                                     (it.shape == SSAM_SHAPE_FACE_EDGE ? 1 : 0)) * 4 + it.storage;
                auto pos = std::find(g3key.begin(), g3key.end(), key);
                if (pos == g3key.end()) {
                    g3key.push_back(key);
                    g3.push_back({i});
                } else {
                    g3[pos - g3key.begin()].push_back(i);
                }
            } else {
                fb.push_back(i);
            }
        } else {
            return SSAM_E_INVALID;
        }
    }
    auto* P = new (std::nothrow) Plan();
    if (!P) return SSAM_E_NOMEM;
    (void)hipGetDevice(&P->device);
    // private copies of the items and their taps (the caller's arrays may go away)
    P->items.assign(items, items + n);
    P->taps_host.resize(n);
    for (int i = 0; i < n; ++i) {
        ssam_batch_item& it = P->items[i];
        const int nt = it.ndim == 2 ? (2 * it.radius + 1) * (2 * it.radius + 1)
                                    : (it.shape == SSAM_SHAPE_FACE_EDGE ? 19 : 6 * it.radius + 1);
        P->taps_host[i].assign(it.coeffs, it.coeffs + nt);
        it.coeffs = P->taps_host[i].data();
        P->point_updates += it.nx * it.ny * it.nz * it.steps;
    }
    P->fallback = fb;
    const int v = env_int("SSAM_BATCH3D_VARIANT", kDefaultVar3D);
    P->var3d = v >= 0 && v < kNumVar3D ? v : kDefaultVar3D;
    P->zseg3d = std::max(1, std::min(env_int("SSAM_BATCH3D_ZSEG", kDefaultZseg3D), 2047));
    const int seg = env_int("SSAM_BATCH2D_SEG", kDefaultSeg2D);
    for (int k = 0; k < kNumKind2D; ++k) P->seg2d[k] = seg >= 8 ? seg : kDefaultSeg2D;
    P->aff2d = env_int("SSAM_BATCH2D_AFFINITY", 1) != 0;
    P->split2d = env_int("SSAM_BATCH2D_SPLIT", 1) != 0;
    P->mfma2d = env_int("SSAM_BATCH2D_MFMA", 3) & 3;
    P->wps2d = std::max(0, env_int("SSAM_BATCH2D_WPS", 0));
    P->wps3d = std::max(0, env_int("SSAM_BATCH3D_WPS", 1));
    P->grid2d = env_int("SSAM_BATCH2D_GRID", 1) != 0;
    P->grid3d = env_int("SSAM_BATCH3D_GRID", 1) != 0;
    P->aff3d = env_int("SSAM_BATCH3D_AFFINITY", 0) != 0;
    P->alt3d = env_int("SSAM_BATCH3D_ALTERNATE", 1) != 0;
    for (int s = 0; s < 4; ++s) {
        if (g2[s].empty()) continue;
        Plan::Group g;
        g.ndim = 2;
        g.storage = s;
        g.ids = g2[s];
        P->groups.push_back(g);
    }
    for (size_t k = 0; k < g3.size(); ++k) {
        Plan::Group g;
        g.ndim = 3;
        const ssam_batch_item& it = items[g3[k][0]];
        g.storage = it.storage;
        g.steps = it.steps;
        g.kind = it.shape == SSAM_SHAPE_FACE_EDGE ? 1 : 0;
        g.ids = g3[k];
        P->groups.push_back(g);
    }
    int rc = build_all(P);
    if (rc == SSAM_OK) {
        // one counter set per 2D launch kind of each 2D group and per (3D group,
        // step) slot of enqueue()
        int n2 = 0, n3 = 0, maxsteps = 0;
        for (auto& g : P->groups) {
            if (g.ndim == 2) ++n2;
            else {
                ++n3;
                maxsteps = std::max(maxsteps, g.steps);
            }
        }
        P->ncounters = std::max(1, kCounterSet * (n2 * kNumLaunch2D + n3 * maxsteps));
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

int ssam_batch_plan_tune(ssam_batch_plan plan, void* stream) {
    if (!plan) return SSAM_E_INVALID;
    Plan* P = plan;
    hipStream_t st = static_cast<hipStream_t>(stream);
    drop_graph(P);  // the tables change; the next run captures anew
    const int reps = 2;
    bool has2 = false, has3 = false, half2 = false, kinds[kNumKind2D] = {false, false, false, false};
    for (auto& g : P->groups) {
        if (g.ndim == 3) {
            has3 = true;
        } else {
            has2 = true;
            half2 = half2 || two_byte(g.storage);
            for (int i : g.ids) {
                const ssam_batch_item& it = P->items[i];
                kinds[kind2d(it.shape == SSAM_SHAPE_STAR ? kStar : kBox, it.radius)] = true;
            }
        }
    }
    int rc = SSAM_OK;
    // Time every combination of candidate values of the given fields (parts of
    // the run only) and keep the fastest.
    auto search = [&](std::vector<int*> fields, std::vector<std::vector<int>> cands, unsigned parts) {
        std::vector<int> best(fields.size()), idx(fields.size(), 0);
        for (size_t f = 0; f < fields.size(); ++f) best[f] = *fields[f];
        float best_ms = 1e30f;
        while (rc == SSAM_OK) {
            for (size_t f = 0; f < fields.size(); ++f) *fields[f] = cands[f][idx[f]];
            float ms = 0.f;
            rc = build_all(P);
            if (rc == SSAM_OK) rc = time_parts(P, st, parts, reps, &ms);
            if (rc == SSAM_OK && ms < best_ms) {
                best_ms = ms;
                for (size_t f = 0; f < fields.size(); ++f) best[f] = *fields[f];
            }
            size_t f = 0;  // next combination (odometer)
            while (f < fields.size() && ++idx[f] == static_cast<int>(cands[f].size())) idx[f++] = 0;
            if (f == fields.size()) break;
        }
        for (size_t f = 0; f < fields.size(); ++f) *fields[f] = best[f];
    };
    if (has2) {
        // per matrix-core routing (the matrix-core tiles want other heights than
        // the VALU ones): launch form x persistent-grid occupancy, then the tile
        // height per shape (one shape at a time, the others held); the faster
        // routing wins; then queue affinity
        struct Cfg { int split, mfma, grid, wps, seg[kNumKind2D]; };
        auto save = [&]() {
            Cfg c{P->split2d, P->mfma2d, P->grid2d, P->wps2d, {}};
            for (int k = 0; k < kNumKind2D; ++k) c.seg[k] = P->seg2d[k];
            return c;
        };
        auto load = [&](const Cfg& c) {
            P->split2d = c.split, P->mfma2d = c.mfma, P->grid2d = c.grid, P->wps2d = c.wps;
            for (int k = 0; k < kNumKind2D; ++k) P->seg2d[k] = c.seg[k];
        };
        const Cfg start = save();
        Cfg best = start;
        float best_ms = 1e30f;
        // (only BF16 / F16 5x5 and 11x11 boxes have a matrix-core form; every
        // subset of the present ones is a routing)
        const int present = half2 ? (kinds[k2Box2] ? 1 : 0) | (kinds[k2Box5] ? 2 : 0) : 0;
        for (int m = 0; m <= 3 && rc == SSAM_OK; ++m) {
            if ((m & ~present) != 0) continue;
            load(start);
            P->mfma2d = present ? m : start.mfma;
            search({&P->split2d, &P->grid2d}, {{0, 1}, {0, 1}}, kRun2D);
            if (!P->grid2d) search({&P->wps2d}, {{1, 2, 4, 0}}, kRun2D);
            for (int k = 0; k < kNumKind2D; ++k)
                if (kinds[k]) search({&P->seg2d[k]}, {{32, 64, 128, 256}}, kRun2D);
            float ms = 0.f;
            if (rc == SSAM_OK) rc = build_all(P);
            if (rc == SSAM_OK) rc = time_parts(P, st, kRun2D, reps, &ms);
            if (rc == SSAM_OK && ms < best_ms) {
                best_ms = ms;
                best = save();
            }
        }
        load(best);
        search({&P->aff2d}, {{0, 1}}, kRun2D);
    }
    if (has3) {
        // tile shape x z-segment length x launch form (persistent / grid), then
        // the persistent grid's occupancy, queue affinity and step direction
        search({&P->var3d, &P->zseg3d, &P->grid3d},
               {{0, 1, 2, 3, kVarPair3D, kVarT2, kVarT3}, {16, 32, 64}, {0, 1}}, kRun3D);
        if (!P->grid3d && kVar3D[P->var3d].t < 2) search({&P->wps3d}, {{1, 2, 4}}, kRun3D);
        search({&P->aff3d, &P->alt3d}, {{0, 1}, {0, 1}}, kRun3D);
    }
    const int rc2 = build_all(P);
    return rc != SSAM_OK ? rc : rc2;
}

int ssam_batch_plan_tiling(ssam_batch_plan plan, int32_t* seg_rows_2d, int32_t* variant_3d,
                           int32_t* zseg_3d, int32_t* sched) {
    if (!plan) return SSAM_E_INVALID;
    if (sched) {
        sched[0] = plan->aff2d;
        sched[1] = plan->aff3d;
        sched[2] = plan->alt3d;
        sched[3] = plan->split2d;
        sched[4] = plan->mfma2d;
        sched[5] = plan->wps2d;
        sched[6] = plan->wps3d;
        sched[7] = plan->grid3d || kVar3D[plan->var3d].t >= 2 ? 1 : 0;
        sched[8] = plan->grid2d;
    }
    if (seg_rows_2d)
        for (int k = 0; k < kNumKind2D; ++k) seg_rows_2d[k] = plan->seg2d[k];
    if (variant_3d) {
        // the shape the plan's 3D groups run (BF16 / F16 groups: always two sweeps)
        *variant_3d = plan->var3d;
        for (const auto& g : plan->groups)
            if (g.ndim == 3) {
                *variant_3d = var3d_of(plan, g);
                if (!two_byte(g.storage)) break;
            }
    }
    if (zseg_3d) *zseg_3d = plan->zseg3d;
    return SSAM_OK;
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
