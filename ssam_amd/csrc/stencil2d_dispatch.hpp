// stencil2d_dispatch.hpp -- host-side launch of the 2D SSAM kernel for one storage
// type: picks the compile-time instance (shape, radius, fused steps), sizes the
// tiles and converts the taps to the accumulation type.
#pragma once
#include "stencil2d.hpp"
#include "stencil2d_lds.hpp"
#include "stencil2d_mfma.hpp"
#include <algorithm>
#include <type_traits>
#include <vector>

#include "ssam_internal.hpp"

namespace ssam {

// Waves the device holds at once for a kernel of 4-wave workgroups: resident
// workgroups per CU (the occupancy API, which accounts for the kernel's VGPRs,
// SGPRs and LDS) x 4 x CUs; cached per kernel by the caller.
inline long long resident_waves(const void* kernel) {
    int dev = 0, cus = 0, blocks = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, kWave * kWavesPerBlock2D, 0) !=
            hipSuccess || blocks < 1)
        return 0;
    return static_cast<long long>(blocks) * kWavesPerBlock2D * cus;
}

// Output rows per tile for a grid of `nstrips` column strips.  The tile count
// aims at waves_per_simd * 1024 waves, rounded to a whole number of device-wide
// rounds of `slots` resident waves (so no round holds a handful of stragglers:
// 4130 tiles on 4096 slots ran a second round for 34 of them), and never below
// min_seg rows per tile (the 2*T*r halo rows per tile stay a small fraction).
// When min_seg caps the segment count below the target, the rounds are chosen
// among those that min_seg still lets fill completely (the shortest tiles then
// only when the target asks for more than that).
inline int tile_rows(int64_t height, int nstrips, int waves_per_simd, long long slots, int min_seg) {
    const long long target = static_cast<long long>(waves_per_simd) * 1024;
    const long long nsegs_max = std::max<long long>(1, height / std::max(1, min_seg));
    long long nsegs;
    if (slots > 0) {
        const long long rounds = std::max<long long>(1, (target + slots / 2) / slots);
        const long long full = nsegs_max * nstrips / slots;  // rounds min_seg lets fill
        // rounded down: at most rounds * slots tiles
        nsegs = rounds <= full ? rounds * slots / nstrips : nsegs_max;
    } else {
        nsegs = target / nstrips;
    }
    nsegs = std::max<long long>(1, std::min(nsegs, nsegs_max));
    long long seg = (height + nsegs - 1) / nsegs;
    if (seg < min_seg) seg = min_seg;
    if (seg > height) seg = height;
    return static_cast<int>(seg);
}

// The ssam_dtype of a storage type.
template <typename S> constexpr int storage_of() {
    return std::is_same<S, float>::value    ? SSAM_F32
           : std::is_same<S, double>::value ? SSAM_F64
           : std::is_same<S, bf16_t>::value ? SSAM_BF16
                                            : SSAM_F16;
}

// Per-storage-type geometry: columns per lane for a single sweep (V: one 16-byte
// access; VN: the narrower strip of 2-byte stars of radius >= 3, boxes of radius
// >= 3 and batch tiles), and for fused time steps of radius 1 / radius 2 (VT1 /
// VT2), and the most steps fused per launch at radius 2.  Fused steps: fewer
// columns keep the T-stage register rings in VGPRs with room for more waves per
// SIMD (FP32 FMAs reach their double rate only when two waves issue together):
// radius 1 at 2 columns (FP32 T=8 10% faster than at 4, BF16 20%); radius 2 up
// to T=4 at 4 columns, which halves the DPP hops per FMA (two T=4 launches of the
// 9-point star on 8192^2: FP32 0.283 -> 0.251 ms, BF16 0.251 -> 0.235).
template <typename S> struct Lanes2D;
template <> struct Lanes2D<float> { static constexpr int V = 4, VN = 4, VT1 = 2, VT2 = 4, TMAX2 = 8; };
template <> struct Lanes2D<double> { static constexpr int V = 2, VN = 2, VT1 = 2, VT2 = 2, TMAX2 = 4; };
template <> struct Lanes2D<bf16_t> { static constexpr int V = 8, VN = 4, VT1 = 2, VT2 = 4, TMAX2 = 8; };
template <> struct Lanes2D<f16_t> { static constexpr int V = 8, VN = 4, VT1 = 2, VT2 = 4, TMAX2 = 8; };


// Loop form per instance.  RING (no register moves): the body is a full ring
// period, lcm(2r+1+PF, 2r+1) ticks, used while that stays under about 2048 FMAs
// of code.  Otherwise blocks of U ticks with a shift (U bounded by about 2048
// FMAs per block).  Also the minimum waves per SIMD the register allocator must
// leave room for.
template <int R, int SHAPE, int V, int T> struct Unroll2D {
    static constexpr int K = 2 * R + 1;
    static constexpr int TAPS = SHAPE == kStar ? 4 * R + 1 : K * K;
    static constexpr int BODY = TAPS * V * T;
    // fused steps: the input ring is two step rings long, so one ring period
    // covers both
    static constexpr int PF_RING = T > 1 ? K : (R <= 2 ? 5 : 3);
    static constexpr int NR0 = K + PF_RING;
    static constexpr int NRS = K;
    static constexpr int PERIOD = T > 1 ? (NR0 / cgcd(NR0, NRS)) * NRS : NR0;
    static constexpr bool RING = PERIOD * BODY <= 2048;
    static constexpr int U0 = 2048 / BODY;
    static constexpr int U = RING ? PERIOD : (U0 < 1 ? 1 : (U0 > 6 ? 6 : U0));
    static constexpr int PF = RING ? PF_RING : (U > 3 ? U : 3);
    static constexpr int WPE = 2;
};

template <typename S, int R, int SHAPE, int V, int T, typename SO = S>
static int launch2d_instance(const Req2D& rq, hipStream_t stream) {
    using A = typename Storage<S>::acc;
    using G = Geom2D<R, V, T>;
    using UR = Unroll2D<R, SHAPE, V, T>;
    constexpr int PF = UR::PF;
    constexpr bool RING = UR::RING;
    constexpr int U = UR::U;
    constexpr int WPE = UR::WPE;

    using KT = KernargTaps<A, R, SHAPE>;
    Taps<A, KT::count> tp{};
    Params2D p;
    p.taps = nullptr;
    if (rq.dev_taps) {
        p.taps = rq.dev_taps;
    } else if constexpr (KT::fits) {
        for (int i = 0; i < KT::n; ++i) tp.w[i] = static_cast<A>(rq.taps[i]);
    } else {
        std::vector<A> hw(KT::n);
        for (int i = 0; i < KT::n; ++i) hw[i] = static_cast<A>(rq.taps[i]);
        const int rc = device_tap_copy(hw.data(), hw.size() * sizeof(A), &p.taps);
        if (rc != SSAM_OK) return rc;
    }

    p.width = static_cast<int>(rq.width);
    p.height = static_cast<int>(rq.height);
    p.pitch = rq.pitch;
    p.pitch_out = rq.pitch_out;
    p.nstrips = static_cast<int>((rq.width + G::OW - 1) / G::OW);
    p.boundary = rq.boundary;
    // one vector access per strip needs only element-aligned rows (any pitch)
    p.vec_ok = (reinterpret_cast<uintptr_t>(rq.src) % sizeof(S) == 0) &&
               (reinterpret_cast<uintptr_t>(rq.dst) % sizeof(SO) == 0);
    // rows per tile: enough tiles for ~target waves per SIMD, but at least a few
    // multiples of the halo so the 2*T*r re-read rows stay a small fraction
    auto kernel = ssam2d_kernel<S, R, SHAPE, V, T, PF, RING, U, WPE, SO>;
    static const long long slots = resident_waves(reinterpret_cast<const void*>(kernel));
    int seg = rq.seg_rows;
    if (seg <= 0) {
        // (8 rows at least: small grids need many short tiles to fill the device)
        const int min_seg = 8 * T * R > 8 ? 8 * T * R : 8;
        seg = tile_rows(rq.height, p.nstrips, rq.waves_per_simd, slots, min_seg);
    }
    if (seg > p.height) seg = p.height;
    p.seg_rows = seg;
    const int nsegs = (p.height + seg - 1) / seg;
    const long long ntiles = static_cast<long long>(p.nstrips) * nsegs;
    if (ntiles > (1LL << 30)) return SSAM_E_UNSUPPORTED;
    p.ntiles = static_cast<int>(ntiles);
    p.nblocks = (p.ntiles + kWavesPerBlock2D - 1) / kWavesPerBlock2D;

    hipLaunchKernelGGL(kernel, dim3(p.nblocks), dim3(kWave * kWavesPerBlock2D), 0, stream, tp,
                       static_cast<const S*>(rq.src), static_cast<SO*>(rq.dst), p);
    return hipGetLastError() == hipSuccess ? SSAM_OK : SSAM_E_HIP;
}

// ---------------------------------------------------------------- LDS-tap form
// Compile-time shape of the LDS-tap kernel per radius (FP32): 4 columns per lane
// up to radius 9 (2 beyond, where the input ring of 2r+1+PF+M-1 rows would not fit
// beside the partial sums), 2 output rows per block, rows prefetched 3 ahead; the
// register budget (waves per SIMD the allocator must leave room for) is 4 up to
// radius 5 and 3 beyond.
template <int R> struct LdsShape {
    static constexpr int V = R <= 9 ? 4 : 2;
    static constexpr int M = 2;
    static constexpr int PF = 3;
    static constexpr int WPE = R <= 5 ? 4 : 3;
};

template <typename S, int R, int SHAPE>
static int launch2d_lds_instance(const Req2D& rq, hipStream_t stream) {
    using A = typename Storage<S>::acc;
    using LS = LdsShape<R>;
    using G = Geom2D<R, LS::V, 1>;
    using KT = KernargTaps<A, R, SHAPE>;
    static_assert(KT::fits, "FP32 taps always fit the kernel-argument segment");
    Taps<A, KT::count> tp{};
    Params2D p;
    p.taps = rq.dev_taps;
    if (!rq.dev_taps)
        for (int i = 0; i < KT::n; ++i) tp.w[i] = static_cast<A>(rq.taps[i]);
    p.width = static_cast<int>(rq.width);
    p.height = static_cast<int>(rq.height);
    p.pitch = rq.pitch;
    p.pitch_out = rq.pitch_out;
    p.nstrips = static_cast<int>((rq.width + G::OW - 1) / G::OW);
    p.boundary = rq.boundary;
    p.vec_ok = (reinterpret_cast<uintptr_t>(rq.src) % sizeof(S) == 0) &&
               (reinterpret_cast<uintptr_t>(rq.dst) % sizeof(S) == 0);
    auto kernel = ssam2d_lds_kernel<S, R, SHAPE, LS::V, LS::M, LS::PF, LS::WPE>;
    static const long long slots = resident_waves(reinterpret_cast<const void*>(kernel));
    int seg = rq.seg_rows > 0 ? rq.seg_rows
                              : tile_rows(rq.height, p.nstrips, rq.waves_per_simd, slots, 8 * R > 8 ? 8 * R : 8);
    if (seg > p.height) seg = p.height;
    p.seg_rows = seg;
    const long long ntiles = static_cast<long long>(p.nstrips) * ((p.height + seg - 1) / seg);
    if (ntiles > (1LL << 30)) return SSAM_E_UNSUPPORTED;
    p.ntiles = static_cast<int>(ntiles);
    p.nblocks = (p.ntiles + kWavesPerBlock2D - 1) / kWavesPerBlock2D;
    hipLaunchKernelGGL(kernel, dim3(p.nblocks), dim3(kWave * kWavesPerBlock2D), 0, stream, tp,
                       static_cast<const S*>(rq.src), static_cast<S*>(rq.dst), p);
    return hipGetLastError() == hipSuccess ? SSAM_OK : SSAM_E_HIP;
}

// FP32 boxes and dense filters of radius 2..14 on the LDS-tap form; each shape
// has its own compile unit (stencil2d_lds_<shape>_f32.hip).
template <typename S, int SHAPE> int dispatch_lds(const Req2D& rq, hipStream_t st);
#ifdef SSAM_2D_LDS_UNIT
template <typename S, int SHAPE> int dispatch_lds(const Req2D& rq, hipStream_t st) {
    switch (rq.radius) {
#define SSAM_R(r) \
    case r:       \
        return launch2d_lds_instance<S, r, SHAPE>(rq, st);
        SSAM_R(2) SSAM_R(3) SSAM_R(4) SSAM_R(5) SSAM_R(6) SSAM_R(7) SSAM_R(8) SSAM_R(9) SSAM_R(10)
        SSAM_R(11) SSAM_R(12) SSAM_R(13) SSAM_R(14)
#undef SSAM_R
        default:
            return SSAM_E_UNSUPPORTED;
    }
}
#endif

// radius is a compile-time instance parameter; these expand the runtime switch.
// Stars go to radius 6, compensated boxes to 9 (19 x 19), boxes and even filters
// to 14: every filter size 2..28 the reference conv2D dispatches
// (ssai-2dconv/conv2D.cu:186-212) embeds in a box of radius <= 14.
template <int SHAPE> constexpr int max_radius() {
    return SHAPE == kStar ? 6 : SHAPE == kBoxKahan ? 9 : 14;
}

template <typename S, int SHAPE, int V, int T>
static int dispatch_radius(const Req2D& rq, hipStream_t st) {
    switch (rq.radius) {
#define SSAM_R(r) \
    case r:       \
        return launch2d_instance<S, r, SHAPE, V, T>(rq, st);
        SSAM_R(1) SSAM_R(2) SSAM_R(3) SSAM_R(4) SSAM_R(5) SSAM_R(6)
#undef SSAM_R
        default:
            break;
    }
    if constexpr (SHAPE != kStar && T == 1) {
        switch (rq.radius) {
#define SSAM_R(r) \
    case r:       \
        return launch2d_instance<S, r, SHAPE, V, T>(rq, st);
            SSAM_R(7) SSAM_R(8) SSAM_R(9)
#undef SSAM_R
            default:
                break;
        }
    }
    if constexpr (SHAPE != kStar && SHAPE != kBoxKahan && T == 1) {
        switch (rq.radius) {
#define SSAM_R(r) \
    case r:       \
        return launch2d_instance<S, r, SHAPE, V, T>(rq, st);
            SSAM_R(10) SSAM_R(11) SSAM_R(12) SSAM_R(13) SSAM_R(14)
#undef SSAM_R
            default:
                break;
        }
    }
    return SSAM_E_UNSUPPORTED;
}

// Fused steps (T in {2, 4, 8}) of a radius-1 or radius-2 star from S to SO; the
// columns per lane are S's (the wider of the two element types decides nothing:
// both sides of a mixed chain use the same tiling).
template <typename S, typename SO>
static int dispatch_fused(const Req2D& rq, hipStream_t st) {
    using LN = Lanes2D<S>;
    constexpr int VT1 = LN::VT1, VT2 = LN::VT2;
    switch (rq.steps) {
        case 1:
            // the single-step tail of a mixed chain (steps not a power-of-two sum of 8s)
            if constexpr (!std::is_same<S, SO>::value) {
                return rq.radius == 1 ? launch2d_instance<S, 1, kStar, VT1, 1, SO>(rq, st)
                                      : launch2d_instance<S, 2, kStar, VT2, 1, SO>(rq, st);
            }
            return SSAM_E_UNSUPPORTED;
        case 2:
            return rq.radius == 1 ? launch2d_instance<S, 1, kStar, VT1, 2, SO>(rq, st)
                                  : launch2d_instance<S, 2, kStar, VT2, 2, SO>(rq, st);
        case 4:
            return rq.radius == 1 ? launch2d_instance<S, 1, kStar, VT1, 4, SO>(rq, st)
                                  : launch2d_instance<S, 2, kStar, VT2, 4, SO>(rq, st);
        case 8:
            if (rq.radius == 1) return launch2d_instance<S, 1, kStar, VT1, 8, SO>(rq, st);
            // eight radius-2 stages on 4 columns per lane run out of registers:
            // 2 columns (the tuner prefers two 4-step launches on 4 columns)
            if constexpr (LN::TMAX2 >= 8) return launch2d_instance<S, 2, kStar, 2, 8, SO>(rq, st);
            return SSAM_E_UNSUPPORTED;
        default:
            return SSAM_E_UNSUPPORTED;
    }
}

// The two halves of a BF16 / FP16 chain of fused launches: 2-byte storage in,
// FP32 running state out (S = bf16/f16), and FP32 in, 2-byte storage out
// (S = float).  Radius-1 and radius-2 stars only.
template <typename S>
static int dispatch_mixed(const Req2D& rq, hipStream_t st) {
    if (rq.shape != kStar || rq.radius > 2) return SSAM_E_UNSUPPORTED;
    if constexpr (sizeof(S) == 2) {
        if (rq.out_storage == SSAM_F32) return dispatch_fused<S, float>(rq, st);
    } else if constexpr (std::is_same<S, float>::value) {
        if (rq.out_storage == SSAM_BF16) return dispatch_fused<float, bf16_t>(rq, st);
        if (rq.out_storage == SSAM_F16) return dispatch_fused<float, f16_t>(rq, st);
    }
    return SSAM_E_UNSUPPORTED;
}

// One sweep of a box-family shape (kBox, kBoxKahan, kBoxEven) on the VALU kernel.
// Each (storage, shape) has its own compile unit (stencil2d_<shape>_<type>.hip,
// SSAM_INSTANTIATE_2D_SHAPE), so the large-radius instances build in parallel.
template <typename S, int SHAPE> int dispatch_shape(const Req2D& rq, hipStream_t st);
#ifdef SSAM_2D_SHAPE_UNIT
template <typename S, int SHAPE> int dispatch_shape(const Req2D& rq, hipStream_t st) {
    // 2-byte storage: 8 columns per lane up to radius 2; larger boxes on the VALU
    // kernel (the matrix-core form takes them by default) use 4, whose ring and
    // resident taps fit the register file
    if constexpr (sizeof(S) == 2) {
        if (rq.radius >= 3) return dispatch_radius<S, SHAPE, Lanes2D<S>::VN, 1>(rq, st);
        if (rq.radius == 1) return launch2d_instance<S, 1, SHAPE, Lanes2D<S>::V, 1>(rq, st);
        return launch2d_instance<S, 2, SHAPE, Lanes2D<S>::V, 1>(rq, st);
    } else {
        return dispatch_radius<S, SHAPE, Lanes2D<S>::V, 1>(rq, st);
    }
}
#endif

template <typename S>
int dispatch2d(const Req2D& rq, hipStream_t st) {
    constexpr int V = Lanes2D<S>::V, VN = Lanes2D<S>::VN;
    if (rq.out_storage != storage_of<S>()) return dispatch_mixed<S>(rq, st);
    if (rq.steps == 1) {
        // matrix cores: by default from radius 2 (where they beat the VALU kernels
        // on MI355X); the autotuner may pick them for radius 1 too
        if (mfma_eligible<S>(rq) &&
            (rq.form == kFormMfma || (rq.form == kFormAuto && rq.radius >= 2)))
            return launch_mfma<S>(rq, st);
        if (rq.shape == kStar) {
            // BF16/FP16 stars of radius >= 3: 4 columns per lane (8-byte rows per
            // lane) keep the ring in the loop-period form with room for more
            // waves; measured 13-36% faster than 8 columns at radius 5 and 6, 8%
            // at radius 3 (8192^2: 0.051 vs 0.055 ms)
            if constexpr (sizeof(S) == 2) {
                switch (rq.radius) {
                    case 1: return launch2d_instance<S, 1, kStar, V, 1>(rq, st);
                    case 2: return launch2d_instance<S, 2, kStar, V, 1>(rq, st);
                    case 3: return launch2d_instance<S, 3, kStar, VN, 1>(rq, st);
                    case 4: return launch2d_instance<S, 4, kStar, VN, 1>(rq, st);
                    case 5: return launch2d_instance<S, 5, kStar, VN, 1>(rq, st);
                    case 6: return launch2d_instance<S, 6, kStar, VN, 1>(rq, st);
                    default: return SSAM_E_UNSUPPORTED;
                }
            } else {
                return dispatch_radius<S, kStar, V, 1>(rq, st);
            }
        }
        // FP32 boxes and dense filters of radius >= 2: the LDS-tap form when asked
        // for, and by default from radius 3 (where the register-cache kernel's
        // resident taps leave too few waves per SIMD for the double FMA rate)
        if constexpr (std::is_same<S, float>::value) {
            if ((rq.shape == kBox || rq.shape == kBoxEven) && rq.radius >= 2 &&
                (rq.form == kFormLds || (rq.form == kFormAuto && rq.radius >= 3)))
                return rq.shape == kBoxEven ? dispatch_lds<S, kBoxEven>(rq, st) : dispatch_lds<S, kBox>(rq, st);
        }
        // FP64 storage accumulates in FP64: compensation changes nothing there
        if constexpr (sizeof(S) != 8) {
            if (rq.shape == kBoxKahan) return dispatch_shape<S, kBoxKahan>(rq, st);
        }
        // even square filters (F32 / F64; 2-byte storage embeds them in a box)
        if constexpr (sizeof(S) >= 4) {
            if (rq.shape == kBoxEven) return dispatch_shape<S, kBoxEven>(rq, st);
        }
        if (rq.shape == kBoxEven) return SSAM_E_UNSUPPORTED;
        return dispatch_shape<S, kBox>(rq, st);
    }
    // fused time steps: star radius 1 and 2 (the reference's temporal 5-point and
    // 9-point star kernels), T in {2, 4, 8}
    if (rq.shape != kStar || rq.radius > 2) return SSAM_E_UNSUPPORTED;
    return dispatch_fused<S, S>(rq, st);
}

}  // namespace ssam

#define SSAM_INSTANTIATE_2D(S) \
    template int ssam::dispatch2d<S>(const ssam::Req2D&, hipStream_t);

#define SSAM_INSTANTIATE_2D_SHAPE(S, SHAPE) \
    template int ssam::dispatch_shape<S, SHAPE>(const ssam::Req2D&, hipStream_t);

#define SSAM_INSTANTIATE_2D_LDS(S, SHAPE) \
    template int ssam::dispatch_lds<S, SHAPE>(const ssam::Req2D&, hipStream_t);
