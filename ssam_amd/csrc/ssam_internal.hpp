// ssam_internal.hpp -- host-side request structs shared by the C-ABI layer
// (api.cpp) and the per-storage-type dispatch units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/ssam.h"

namespace ssam {

// 2D kernel shapes.  kBoxKahan: box / dense taps with compensated accumulation
// (each filter row is summed on its own, then added to the running sum with
// Kahan compensation).
// kBoxEven: an even square filter (2r x 2r taps, stored densely) in the
// (2r+1)^2 box, whose last row and column are zero.
enum Shape2D : int { kStar = 0, kBox = 1, kBoxKahan = 2, kBoxEven = 3 };

// A validated 2D request: one launch of `steps` fused sweeps (steps in 1,2,4,8).
struct Req2D {
    const void* src;
    void* dst;
    int64_t width, height, pitch;
    // dst's storage (ssam_dtype; the request's storage when equal) and row pitch:
    // FP32 for the running state between the fused launches of a BF16 / FP16 chain
    int out_storage;
    int64_t pitch_out;
    int shape;       // Shape2D (DENSE arrives embedded in a box)
    int radius;      // filter half-width after embedding
    int steps;       // fused steps of this launch
    int boundary;    // Boundary
    const double* taps;  // host taps, row-major: (2r+1)^2, or (2r)^2 for kBoxEven;
                         // nullptr when dev_taps is set
    // Device taps in the accumulation type, read in place by the kernel (the
    // reference conv2D's device weights), laid out as the shape's host taps.
    const void* dev_taps;
    int seg_rows;        // 0 = heuristic
    int waves_per_simd;  // tile-count target for the heuristic
    int form;            // Form2D: kernel family (auto picks the matrix cores where they apply)
};

// 2D kernel family: register-cache VALU kernels, the matrix-core (MFMA) form, or
// the LDS-tap VALU form (FP32 boxes and dense filters; taps staged in LDS).
enum Form2D : int { kFormAuto = 0, kFormValu = 1, kFormMfma = 2, kFormLds = 3 };

template <typename S> int dispatch2d(const Req2D& rq, hipStream_t st);

// Taps too large for the kernel-argument segment (FP64 filters past 22 x 22
// given as host coefficients): a device copy in the accumulation type, kept for
// the life of the process and keyed by its contents, so only the first call with
// a given filter uploads (synchronously); SSAM_OK or an error status.
int device_tap_copy(const void* host, size_t bytes, const void** dev);

// A validated 3D request: one sweep.
struct Req3D {
    const void* src;
    void* dst;
    int out_storage;     // ssam_dtype of dst (the request's storage, or FP32 / 2-byte
                         // at the ends of a chain with FP32 running state)
    int64_t nx, ny, nz;
    int shape;           // Shape3D: 0 star, 1 face+edge (19 points)
    int radius;
    const double* taps;  // descriptor order (include/ssam.h)
    int zseg;            // output planes per tile, 0 = heuristic
    int form;            // radius-2 stars: 0 auto, 1 plane ring, 2 z-partial sums
    int sweeps;          // sweeps per launch of dispatch3d_two_sweeps: 2, or 3 (7-point star)
    double* partials;    // device, per-tile residual partials (nullptr: none)
    int partials_cap;
    double* res_out;     // device, the summed residual^2 (or its root, res_sqrt)
    int res_sqrt;
    double res_scale;
};

// One sweep from `storage` (src) to rq.out_storage (dst).
int dispatch3d(int storage, const Req3D& rq, hipStream_t st);
// Two sweeps in one launch (radius 1, grid under 2 GiB), or three (rq.sweeps == 3,
// the 7-point star); SSAM_E_UNSUPPORTED else.
bool two_sweeps_ok(int storage, const Req3D& rq);
int dispatch3d_two_sweeps(int storage, const Req3D& rq, hipStream_t st);
int partials_needed(int storage, const Req3D& rq);

}  // namespace ssam
