/*
 * ssam.h -- C ABI of libssam.so, the MI355X (gfx950) SSAM stencil / convolution
 * library.
 *
 * Conventions (all entry points):
 *   - device pointers are caller-owned HIP device memory; the library never
 *     allocates grids and never aliases src and dst (overlap -> SSAM_E_INVALID);
 *   - sizes and pitches are in elements; pitch >= width;
 *   - `stream` is a hipStream_t passed as void* (NULL = the default stream); every
 *     call is asynchronous on that stream unless its comment says otherwise;
 *   - the return value is an ssam_status; the library never aborts or throws.
 * The reference (peng-codes/SSAM) has no library API: its callers are the C++
 * harness entry points `int stencil_<pts>_<dtype>(int argc, char** argv)`
 * (ssai-2dstencil/main.cpp:29-53, ssai-3dstencil/main.cpp:5-9) launching the
 * templated kernels cited per entry below.  The drivers under
 * ssam_amd/drivers/ keep those command lines and call these functions.
 */
#ifndef SSAM_H
#define SSAM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
typedef enum {
    SSAM_OK = 0,
    SSAM_E_INVALID = 1,     /* bad argument: null pointer, size, shape, aliasing */
    SSAM_E_UNSUPPORTED = 2, /* well-formed request this build has no kernel for */
    SSAM_E_HIP = 3,         /* HIP runtime reported an error (launch, device) */
    SSAM_E_NOMEM = 4        /* host-side staging allocation failed */
} ssam_status;

/* ---------------------------------------------------------------- enums */
typedef enum {
    SSAM_F32 = 0,  /* FP32 storage, FP32 accumulation */
    SSAM_F64 = 1,  /* FP64 storage, FP64 accumulation */
    SSAM_BF16 = 2, /* BF16 storage, FP32 accumulation and running state */
    SSAM_F16 = 3   /* FP16 storage, FP32 accumulation and running state */
} ssam_dtype;

typedef enum {
    SSAM_SHAPE_STAR = 0,     /* 2D: centre row + centre column of a (2r+1)^2 box;
                                3D: the 6r+1-point axis star */
    SSAM_SHAPE_BOX = 1,      /* 2D: full (2r+1)^2 box */
    SSAM_SHAPE_DENSE = 2,    /* 2D: any fw x fh filter (even sizes allowed) */
    SSAM_SHAPE_FACE_EDGE = 3 /* 3D: 19 points = centre, 6 faces, 12 edges */
} ssam_shape;

typedef enum {
    SSAM_BOUNDARY_ZERO = 0,      /* reads outside the grid are 0 (reference CPU gold,
                                    ssai-2dconv/common.h:200-206, and conv2D.cu:36-41) */
    SSAM_BOUNDARY_CLAMP = 1,     /* reads clamp to the edge (reference GPU stencils,
                                    stencil2D_5pt-v1.cu:29-46); steps == 1 only */
    SSAM_BOUNDARY_HOLD = 2,      /* cells within r of the border keep their input
                                    value; steps == 1 only */
    SSAM_BOUNDARY_DIRICHLET0 = 3 /* 3D: the outer r layers of every output are 0
                                    (reference 3D gold, tmp3d7pt-v2.cu:284-296) */
} ssam_boundary;

#define SSAM_FLAG_WORK_FP32 2u   /* BF16 / FP16 storage: `work` holds 2 * width *
                                    height FP32 values (dense rows), where a run of
                                    several fused launches keeps its running state
                                    in FP32 between launches.  Without it such a
                                    run is SSAM_E_INVALID (the state is never
                                    rounded to the storage type). */
#define SSAM_FLAG_COMPENSATED 1u /* Kahan-compensated accumulation: each filter
                                    row is summed on its own and added to the
                                    running sum with compensation.  BOX and DENSE
                                    with steps == 1 (else SSAM_E_UNSUPPORTED);
                                    FP64 storage accumulates in FP64 already and
                                    is unchanged by it. */

/* ---------------------------------------------------------------- descriptors */
typedef struct {
    int32_t storage;      /* ssam_dtype of src, dst and work */
    int32_t shape;        /* STAR, BOX or DENSE */
    int32_t fw, fh;       /* filter width / height; STAR and BOX: fw == fh == 2r+1;
                             STAR r <= 6, BOX r <= 14, DENSE up to 29 x 29
                             (COMPENSATED: up to 19 x 19) */
    int32_t steps;        /* time steps applied (>= 1); > 1 only for ZERO boundary */
    int32_t boundary;     /* ssam_boundary: ZERO, CLAMP or HOLD */
    uint32_t flags;       /* SSAM_FLAG_COMPENSATED | SSAM_FLAG_WORK_FP32 (other bits:
                             SSAM_E_INVALID) */
    const double* coeffs; /* HOST memory, fh*fw taps row-major: w[m*fw + n] weights
                             in(x + n - fw/2, y + m - fh/2) (correlation, no flip).
                             STAR reads only the centre row and column. */
} ssam_stencil2d_desc;

typedef struct {
    int32_t storage;      /* ssam_dtype: F32, F64, BF16 or F16 */
    int32_t shape;        /* STAR (radius 1 or 2) or FACE_EDGE */
    int32_t radius;       /* 1 or 2 for STAR, 1 for FACE_EDGE */
    int32_t steps;        /* sweeps (>= 1) */
    int32_t boundary;     /* SSAM_BOUNDARY_DIRICHLET0 */
    uint32_t flags;       /* SSAM_FLAG_WORK_FP32 (BF16 / F16 with steps > 1: `work`
                             holds 2 * nx*ny*nz FP32 values); other bits:
                             SSAM_E_INVALID */
    const double* coeffs; /* HOST memory.  STAR r: 6r+1 taps ordered
                             c, x-1, x+1, y-1, y+1, z-1, z+1, x-2, x+2, y-2, y+2, z-2, z+2.
                             FACE_EDGE: 19 taps ordered c, 6 faces (as STAR r=1), then
                             edges (y-1,z-1) (y+1,z-1) (y-1,z+1) (y+1,z+1)
                             (x-1,z-1) (x+1,z-1) (x-1,z+1) (x+1,z+1)
                             (x-1,y-1) (x+1,y-1) (x-1,y+1) (x+1,y+1). */
} ssam_stencil3d_desc;

/* ---------------------------------------------------------------- library */
const char* ssam_version(void);
const char* ssam_strerror(int status);

/* ---------------------------------------------------------------- 2D generic
 * Applies `steps` sweeps of the 2D stencil / correlation described by `desc` to
 * src (width x height, row pitch `pitch`), result in dst.  Temporal steps are
 * fused in registers up to 8 per launch; a `steps` needing more than one launch
 * ping-pongs through `work` (may be NULL otherwise): F32 / F64 storage -- the
 * same shape as dst; BF16 / F16 storage -- 2 * width * height FP32 values, with
 * SSAM_FLAG_WORK_FP32 set (the running state stays FP32 between launches).
 * Replaces the kernel launches of stencil2D_5pt-v1.cu:194-195, stencil2D_*pt.cu,
 * stencil2d_temp_{5,9}pt.cu:146-147 and conv2D.cu:128-131. */
int ssam_stencil2d(const ssam_stencil2d_desc* desc, const void* src, void* dst, int64_t width,
                   int64_t height, int64_t pitch, void* work, void* stream);

/* ---------------------------------------------------------------- 2D, reference kernels
 * stencil2d_v1::j2d5pt(src, dst, width, height, fc, fn, fs, fw, fe)
 * (ssai-2dstencil/stencil2D_5pt-v1.cu:8-9); zero boundary, pitch = width. */
int ssam_j2d5pt_f32(const float* src, float* dst, int width, int height, float fc, float fn,
                    float fs, float fw, float fe, void* stream);
int ssam_j2d5pt_f64(const double* src, double* dst, int width, int height, double fc, double fn,
                    double fs, double fw, double fe, void* stream);

/* Star of radius r = 1..6 (j2d9pt .. j2d25pt, stencil2D_{9,13,17,21,25}pt.cu:9-11).
 * coefs[4r+1] in the reference kernels' argument order:
 *   fc, fn0, fs0, fw0, fe0, fn1, fs1, fw1, fe1, ...  (fnk = k+1 rows north). */
int ssam_j2dstar_f32(const float* src, float* dst, int width, int height, int radius,
                     const float* coefs, void* stream);
int ssam_j2dstar_f64(const double* src, double* dst, int width, int height, int radius,
                     const double* coefs, void* stream);

/* Box of radius r = 1..14 (j2d64pt / j2d81pt / j2d121pt, stencil2D_64pt.cu:74,
 * stencil2D_81pt.cu:92, stencil2D_121pt.cu:132, which bake their taps in as
 * macros; here they are a host array w[(2r+1)^2], row-major). */
int ssam_j2dbox_f32(const float* src, float* dst, int width, int height, int radius,
                    const float* w, void* stream);
int ssam_j2dbox_f64(const double* src, double* dst, int width, int height, int radius,
                    const double* w, void* stream);

/* Temporally blocked 5-point (stencil2d_tmp::j2d5pt, stencil2d_temp_5pt.cu:9-10) and
 * 9-point star (stencil2d_tmp::j2d9pt, stencil2d_temp_9pt.cu:12-14): `steps`
 * zero-padded sweeps; `work` as for ssam_stencil2d. */
int ssam_j2d5pt_temporal_f32(const float* src, float* dst, int width, int height, float fc,
                             float fn, float fs, float fw, float fe, int steps, float* work,
                             void* stream);
int ssam_j2d5pt_temporal_f64(const double* src, double* dst, int width, int height, double fc,
                             double fn, double fs, double fw, double fe, int steps,
                             double* work, void* stream);
int ssam_j2d9pt_temporal_f32(const float* src, float* dst, int width, int height, float fc,
                             float fn0, float fs0, float fw0, float fe0, float fn1, float fs1,
                             float fw1, float fe1, int steps, float* work, void* stream);
int ssam_j2d9pt_temporal_f64(const double* src, double* dst, int width, int height, double fc,
                             double fn0, double fs0, double fw0, double fe0, double fn1,
                             double fs1, double fw1, double fe1, int steps, double* work,
                             void* stream);

/* CONV2D::kernel_convolution2D(src, dst, width, widthStride, height, weight)
 * (ssai-2dconv/conv2D.cu:9-12): FW x FH correlation, zero padding, pitched rows,
 * filter sizes up to 28 x 28 (the reference dispatches 2..28, conv2D.cu:186-212;
 * 29 x 29 is accepted too).  `weight` is DEVICE memory (fh*fw, row-major) as in the
 * reference.  Square filters (fw == fh >= 2) are read by the kernel in place: the
 * call is asynchronous and graph-capturable, and `weight` must stay valid until
 * the stream has run it.  Other shapes (and 1 x 1) are read back once per call
 * and embedded on the host, so those calls SYNCHRONISE `stream`. */
int ssam_conv2d_f32(const float* src, float* dst, int width, int width_stride, int height,
                    const float* weight, int fw, int fh, void* stream);
int ssam_conv2d_f64(const double* src, double* dst, int width, int width_stride, int height,
                    const double* weight, int fw, int fh, void* stream);

/* ---------------------------------------------------------------- 3D
 * `steps` sweeps of a 3D stencil on a dense nx*ny*nz grid (index z*ny*nx + y*nx + x,
 * tmp3d7pt-v2.cu:107).  Every output point is written: the outer r layers are 0
 * (SSAM_BOUNDARY_DIRICHLET0), the interior gets the stencil of the input.
 * steps > 1 ping-pongs through `work` (may be NULL otherwise): F32 / F64 -- the
 * grid's size; BF16 / F16 storage (FP32 accumulation) -- 2 * nx*ny*nz FP32 values
 * with SSAM_FLAG_WORK_FP32 set, where the running state stays FP32 between
 * launches (the grid is read and written in the storage type once each).
 * Replaces kernel3d_3d7pt (tmp3d7pt-v2.cu:62-64), kernel3d_3d13pt
 * (tmp3d13pt.cu:74-76), kernel3d_poisson (tmpPoisson.cu:86-88). */
int ssam_stencil3d(const ssam_stencil3d_desc* desc, const void* src, void* dst, int64_t nx,
                   int64_t ny, int64_t nz, void* work, void* stream);

/* Reference operators with their reference coefficients (tmp3d7pt-v2.cu:9-15,
 * tmp3d13pt.cu:9-21, tmpPoisson.cu:9-27), `iters` sweeps. */
int ssam_j3d7pt_f32(const float* src, float* dst, int n, int iters, float* work, void* stream);
int ssam_j3d7pt_f64(const double* src, double* dst, int n, int iters, double* work, void* stream);
int ssam_j3d13pt_f32(const float* src, float* dst, int n, int iters, float* work, void* stream);
int ssam_j3d13pt_f64(const double* src, double* dst, int n, int iters, double* work,
                     void* stream);
int ssam_poisson_f32(const float* src, float* dst, int n, int iters, float* work, void* stream);
int ssam_poisson_f64(const double* src, double* dst, int n, int iters, double* work,
                     void* stream);

/* 19-point Jacobi sweeps for the 3D Laplace problem (this library's operator; the
 * reference's "poisson" grows without bound when iterated):
 *   u'(c) = (2*sum_faces u + sum_edges u) / 24, zero Dirichlet boundary.
 * F32 / F64 storage: running state in the storage type, `work` the grid's size;
 * BF16 / F16 storage: the running state stays FP32 and `work` holds
 * 2 * nx*ny*nz FP32 values (steps > 1).  If residual_l2 is
 * not NULL it receives the FP64 L2 norm of r = -A u of the input of the LAST
 * sweep, A u = (24 u - 2 faces - edges)/6, reduced on the GPU in a fixed order;
 * this entry then synchronises `stream`. */
int ssam_jacobi19(int storage, const void* src, void* dst, int64_t nx, int64_t ny, int64_t nz,
                  int steps, void* work, double* residual_l2, void* stream);

/* ssam_jacobi19 for callers that cannot allocate (code under hipGraph stream
 * capture, tight serving loops): no counterpart in the reference.  Same sweeps
 * and bit-identical results, but it allocates nothing -- the residual's per-tile
 * partials go to a fixed 1 MiB device array inside the library (one per GPU,
 * held until the process ends), grids needing more than 131072 partials are
 * SSAM_E_UNSUPPORTED -- and `residual_l2` is a DEVICE double (may be NULL)
 * receiving the FP64 L2 norm, so the call is asynchronous.  Calls that overlap
 * on one GPU share that array and race; the first call on a device looks up the
 * array's address, so make it outside graph capture. */
int ssam_jacobi19_noalloc(int storage, const void* src, void* dst, int64_t nx, int64_t ny,
                          int64_t nz, int steps, void* work, double* residual_l2, void* stream);

/* ---------------------------------------------------------------- batches
 * Many independent grids of uneven sizes and mixed shapes.  2D items of one
 * storage type whose shape is a 5-point star or a 3x3 / 5x5 / 11x11 box run as
 * one launch per stencil kind (or one for all kinds); 3D items (7-point star or
 * 19-point face+edge, any storage, under 2 GiB) run one launch per pass of one,
 * two or three sweeps for each (storage, steps, shape) -- BF16 / F16 items two or
 * three sweeps per pass with their state FP32 between passes; any other item gets
 * its own launch.
 * Each launch is an ordinary grid with one tile per wave over a table sorted by
 * decreasing tile cost (default), or a persistent grid taking tiles from 8 device
 * work queues, one per XCD (neighbouring tiles of a grid share that XCD's L2),
 * stealing from the other queues when its own is empty; the tuner picks.  The
 * first ssam_batch_run captures the whole sequence into a hipGraph on a private
 * stream and every run replays it on `stream` (SSAM_BATCH_EAGER: launch
 * directly).  Device pointers are bound at plan creation; items and taps are
 * copied.  No counterpart in the
 * reference. */
typedef struct {
    int32_t ndim;          /* 2 or 3 */
    int32_t storage;       /* ssam_dtype */
    int32_t shape;         /* 2D: STAR or BOX; 3D: STAR or FACE_EDGE */
    int32_t radius;        /* filter radius r */
    int32_t steps;         /* 2D: 1 in the persistent launch; 3D: sweeps */
    int32_t reserved;
    int64_t nx, ny, nz;    /* 2D: width, height, 1 */
    int64_t pitch;         /* 2D row pitch in elements; 3D grids are dense */
    const void* src;
    void* dst;
    void* work;            /* 3D with steps > 1: ping-pong buffer of the same size
                              (BF16 / F16: 2 * nx*ny*nz FP32 values, the state) */
    const double* coeffs;  /* HOST: 2D (2r+1)^2 row-major; 3D descriptor tap order */
} ssam_batch_item;

typedef struct ssam_batch_plan_s* ssam_batch_plan;
#define SSAM_BATCH_EAGER 1u

int ssam_batch_plan_create(const ssam_batch_item* items, int n, ssam_batch_plan* plan);
int ssam_batch_run(ssam_batch_plan plan, void* stream, uint32_t flags);
int ssam_batch_plan_info(ssam_batch_plan plan, int64_t* tiles_2d, int64_t* tiles_3d,
                         int64_t* point_updates);
/* Autotune the plan's tiling on its own buffers (dst and work are overwritten;
 * the stream is synchronised): the 2D launch forms (one launch per stencil kind
 * or one for all; persistent or grid) and whether BF16/FP16 5x5 and 11x11 boxes
 * run on the matrix cores, the 2D tile height per shape (32..256 rows), the 3D
 * tile shape (columns and rows per lane, prefetch depth), z-segment length and
 * launch form, the persistent grid's waves per SIMD and the queue affinity of
 * each, and the 3D sweep direction, timed over eager runs; the fastest is kept
 * and the next ssam_batch_run captures a new graph. */
int ssam_batch_plan_tune(ssam_batch_plan plan, void* stream);
/* The plan's tiling: 2D tile height per shape (4 entries: 5-point star, 3x3,
 * 5x5, 11x11 box), 3D tile-shape index (0..3 one sweep per pass; 4: per grid,
 * the better of shapes 2 and 3 by width; 5: two sweeps per pass; 6: three sweeps
 * per pass for 7-point items, two for 19-point ones) and z-segment length, and 9
 * scheduling
 * entries: queue affinity of the 2D and of the 3D launches (1: waves start at
 * their own XCD's queue, 0: all waves work through the tiles in table order),
 * whether 3D sweeps alternate direction (odd steps walk the table back to front),
 * the 2D launch form (1: one persistent launch per stencil kind, 0: one for all
 * kinds), the matrix-core routing of BF16/FP16 boxes (bit 0: 5x5, bit 1: 11x11
 * boxes on the matrix cores), the waves per
 * SIMD of the 2D and of the 3D persistent grids (0: as many as fit), and the 3D
 * and 2D launch forms (1: an ordinary grid with one tile per wave, the hardware's
 * workgroup dispatch taking the place of the queues; 0: persistent).  Any
 * pointer may be NULL. */
int ssam_batch_plan_tiling(ssam_batch_plan plan, int32_t* seg_rows_2d, int32_t* variant_3d,
                           int32_t* zseg_3d, int32_t* sched);
void ssam_batch_plan_destroy(ssam_batch_plan plan);

/* ---------------------------------------------------------------- micro-benchmarks
 * CDNA4 measurements behind the kernels' design (counterpart of the reference's
 * shuffle / shared-memory / FMA latency measurements, SSAM paper Table 2):
 *   which 0..3: latency in cycles of one dependent operation on one wave --
 *     0 FP32 FMA, 1 DPP wave_shr:1 fused into v_add_f32 (the systolic hop),
 *     2 ds_bpermute + add, 3 LDS store + load round trip + add;
 *   which 4..8: wave-instructions per cycle per CU over the whole device, 8 waves
 *     per CU with 8 independent chains each -- 4 FMA, 5 DPP add, 6 ds_bpermute,
 *     7 LDS round trip, 8 v_mfma_f32_16x16x32_bf16;
 *   which 9..13: device rate in TFLOP/s, 32 waves per CU with 8 independent
 *     chains, timed with HIP events (wall clock) -- 9 FP32 FMA (no packed FP32),
 *     10 FP64 FMA, 11 v_mfma_f32_16x16x4_f32, 12 v_mfma_f32_4x4x1_16b_f32,
 *     13 v_mfma_f64_16x16x4_f64;
 *   which 14..19: FP32 FMA operand forms, TFLOP/s of the FMAs -- 14 every operand
 *     a VGPR, 15 one SGPR tap, 16 sixteen kernel-argument taps, 17 SGPR taps with
 *     one DPP add per 4 FMAs, 18 VGPR taps with 2 DPP adds per 10 FMAs (the hop's
 *     issue cost), 19 the same with 2 plain adds.
 * Cycles (0..8) are s_memtime counts.  Synchronises `stream`. */
int ssam_microbench(int which, double* result, void* stream);

/* ---------------------------------------------------------------- autotuner
 * Times candidate tilings of one 2D request on the caller's buffers (dst and work
 * are overwritten; the stream is synchronised): tile height, as the number of
 * waves per SIMD the tiling targets (2, 4, 8, 12), and the fused steps per launch
 * (temporal-blocking degree, 8/4/2/1 where the shape fuses and `work` is given).
 * The best form is cached in the process, keyed by (storage, shape, radius,
 * steps, width, height), and used by every later ssam_stencil2d call with that
 * key.  best_* (may be NULL) receive the choice.  No counterpart in the reference,
 * whose tile parameters are compile-time tables keyed by GPU name
 * (tmp3d7pt-v2.cu:341-431). */
int ssam_tune_stencil2d(const ssam_stencil2d_desc* desc, const void* src, void* dst,
                        int64_t width, int64_t height, int64_t pitch, void* work, void* stream,
                        int* best_waves_per_simd, int* best_tchunk);
void ssam_tune_clear(void);
/* Persist / restore the cache as text lines; load returns SSAM_E_INVALID if the
 * file holds no valid entry. */
int ssam_tune_save(const char* path);
int ssam_tune_load(const char* path);

/* ---------------------------------------------------------------- utilities
 * Seeded synthetic input: element i = a + (b-a) * (splitmix64_{seed}(offset+i) >> 40)
 * * 2^-24, written in `dtype`.  Same values as the CPU generator of the oracle. */
int ssam_fill_uniform(void* dst, int dtype, int64_t n, uint64_t seed, uint64_t offset, double a,
                      double b, void* stream);

/* Device copy of `bytes` (multiple of 16, 16-byte aligned) with dwordx4 loads and
 * stores: the measured HBM stream bandwidth the roofline fractions refer to. */
int ssam_stream_copy(const void* src, void* dst, int64_t bytes, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* SSAM_H */
