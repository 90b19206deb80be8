/*
 * oracle/gold.c -- CPU restatement of the SSAM reference's gold loops.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load the
 * library built from it.  The product (ssam_amd/libssam.so) never links it.
 *
 * Build: oracle/Makefile  ->  oracle/liboracle.so
 *   gcc -O3 -fopenmp -ffp-contract=off   (the reference gold was built without
 *   FMA contraction, ssai-2dstencil/Makefile:67 `-msse4.1`, so products and sums
 *   round separately here too).
 *
 * Each function cites the reference lines it restates.  Loop order and the
 * accumulation type follow the reference, so FP32 results are bit-identical to
 * the reference gold compiled the same way (pinned against oracle/_ref, which is
 * built from the reference's own common.h, in tests/test_oracle.py).
 *
 * Parity pins (tests/test_oracle.py):
 *   orc_conv2d_*         : oracle/_ref (reference common.h:186-216 compiled as-is),
 *                          impulse response, constant field, brute force in numpy.
 *   orc_conv2d_iter_*    : equals `steps` single sweeps (closed form), _ref iterated.
 *   orc_stencil3d_taps_* : impulse response against the reference coefficient
 *                          fixtures (tests/golden/coeffs.json), constant field.
 *   orc_poisson_ref_*    : impulse response + constant field against the literals
 *                          of tmpPoisson.cu:378-382.
 *   orc_jacobi19_*       : convex-combination invariants (constant field fixed
 *                          point, maximum principle), impulse response.
 *   orc_uniform_*        : splitmix64 published test vector (seed 1234567).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------- */
/* Seeded generator: splitmix64 (Steele, Lea, Flood 2014; Vigna's reference   */
/* C).  Element i is the (i+1)-th output of a splitmix64 stream whose state   */
/* starts at `seed`, so the stream is random-access.  Replaces the reference's */
/* glibc rand() (common.h:14-17, common.hpp:56-74), which explicitly          */
/* seeded, portable inputs cannot reproduce.                                 */
/* ------------------------------------------------------------------------- */
static inline uint64_t orc_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

ORC_API uint64_t orc_splitmix64_at(uint64_t seed, uint64_t i) {
    return orc_mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ULL);
}

/* value = a + (b - a) * (u >> 40) * 2^-24 ; exact in FP32 for [0,1) and [-1,1) */
static inline double orc_u01(uint64_t u) { return (double)(u >> 40) * (1.0 / 16777216.0); }

ORC_API void orc_uniform_f64(uint64_t seed, uint64_t offset, int64_t n, double a, double b,
                             double* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        out[i] = a + (b - a) * orc_u01(orc_splitmix64_at(seed, offset + (uint64_t)i));
}

ORC_API void orc_uniform_f32(uint64_t seed, uint64_t offset, int64_t n, double a, double b,
                             float* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        out[i] = (float)(a + (b - a) * orc_u01(orc_splitmix64_at(seed, offset + (uint64_t)i)));
}

/* ------------------------------------------------------------------------- */
/* 2D gold: dense FW x FH correlation with zero padding.                      */
/* Restates Convolution<T> (ssai-2dconv/common.h:186-216): OpenMP over rows,  */
/* filter column n outer, filter row m inner, taps outside the image skipped, */
/* accumulation in T, every pixel written.  No filter flip (correlation).     */
/* ------------------------------------------------------------------------- */
#define ORC_CONV2D(T, NAME)                                                              \
    ORC_API void NAME(const T* src, T* dst, int width, int height, const T* filter,     \
                      int flt_width, int flt_height) {                                  \
        _Pragma("omp parallel for schedule(static)")                                   \
        for (int i = 0; i < height; i++) {                                              \
            for (int j = 0; j < width; j++) {                                           \
                T sum = 0;                                                              \
                for (int n = 0; n < flt_width; n++) {                                   \
                    for (int m = 0; m < flt_height; m++) {                              \
                        int x = j + n - flt_width / 2;                                  \
                        int y = i + m - flt_height / 2;                                 \
                        if (x >= 0 && x < width && y >= 0 && y < height)                \
                            sum += src[(int64_t)y * width + x] * filter[m * flt_width + n]; \
                    }                                                                   \
                }                                                                       \
                dst[(int64_t)i * width + j] = sum;                                      \
            }                                                                           \
        }                                                                               \
    }
ORC_CONV2D(float, orc_conv2d_f32)
ORC_CONV2D(double, orc_conv2d_f64)

/* Temporal gold: `steps` zero-padded sweeps, the output copied back to the input
 * after each (ssai-2dstencil/stencil2d_temp_5pt.cu:169-175,
 * stencil2d_temp_9pt.cu:180-186).  `img` holds the input on entry and the result
 * on return; `tmp` is a scratch image of the same size. */
#define ORC_CONV2D_ITER(T, NAME, CONV)                                                   \
    ORC_API void NAME(T* img, T* tmp, int width, int height, const T* filter, int fw,   \
                      int fh, int steps) {                                              \
        for (int s = 0; s < steps; s++) {                                               \
            CONV(img, tmp, width, height, filter, fw, fh);                              \
            memcpy(img, tmp, sizeof(T) * (size_t)width * (size_t)height);               \
        }                                                                               \
    }
ORC_CONV2D_ITER(float, orc_conv2d_iter_f32, orc_conv2d_f32)
ORC_CONV2D_ITER(double, orc_conv2d_iter_f64, orc_conv2d_f64)

/* ------------------------------------------------------------------------- */
/* 3D golds.  Layout index z*ny*nx + y*nx + x (tmp3d7pt-v2.cu:107).           */
/* Only the interior [r, n-r) of each axis is written; the caller's output    */
/* boundary is left as it was (zero-initialised in the reference,            */
/* common.hpp:78-82), i.e. zero Dirichlet from the second step on.           */
/*                                                                            */
/* orc_stencil3d_taps_*: out = c0*in[t0] + c1*in[t1] + ... evaluated left to  */
/* right in T.  With the reference tap order and coefficients this restates  */
/*   j3d7pt_gold  (tmp3d7pt-v2.cu:278-298; order z-1, y-1, c, y+1, x-1, x+1, z+1) */
/*   j3d13pt_gold (tmp3d13pt.cu:301-327; order z-2, z-1, c, z+1, z+2, y-1, y-2,  */
/*                 y+1, y+2, x-2, x-1, x+1, x+2)                                 */
/* The coefficients are cast to T first, as the reference's ((T)0.1) macros do. */
/* taps: ntaps triples (dz, dy, dx); radius: interior margin.                  */
/* ------------------------------------------------------------------------- */
#define ORC_TAPS3D(T, NAME)                                                              \
    ORC_API void NAME(const T* in, T* out, int nx, int ny, int nz, int radius,          \
                      int ntaps, const int* taps, const double* coef) {                 \
        const int64_t sy = nx, sz = (int64_t)nx * ny;                                   \
        T c[64];                                                                        \
        int64_t off[64];                                                                \
        if (ntaps > 64) return;                                                         \
        for (int t = 0; t < ntaps; t++) {                                               \
            c[t] = (T)coef[t];                                                          \
            off[t] = taps[3 * t] * sz + taps[3 * t + 1] * sy + taps[3 * t + 2];         \
        }                                                                               \
        _Pragma("omp parallel for schedule(static)")                                   \
        for (int k = radius; k < nz - radius; k++)                                      \
            for (int j = radius; j < ny - radius; j++)                                  \
                for (int i = radius; i < nx - radius; i++) {                            \
                    const int64_t p = k * sz + j * sy + i;                              \
                    T acc = c[0] * in[p + off[0]];                                      \
                    for (int t = 1; t < ntaps; t++) acc = acc + c[t] * in[p + off[t]];  \
                    out[p] = acc;                                                       \
                }                                                                       \
    }
ORC_TAPS3D(float, orc_stencil3d_taps_f32)
ORC_TAPS3D(double, orc_stencil3d_taps_f64)

/* poisson_gold (tmpPoisson.cu:369-386).  The reference writes the literals
 * 2.666 / 0.166 / 0.0833 as double constants, so for T=float the face sum and the
 * signed edge sum are formed in float (left to right, as written) and the final
 * combination is evaluated in double, then stored to T. */
#define ORC_POISSON(T, NAME)                                                             \
    ORC_API void NAME(const T* in, T* out, int nx, int ny, int nz) {                    \
        const int64_t sy = nx, sz = (int64_t)nx * ny;                                   \
        _Pragma("omp parallel for schedule(static)")                                   \
        for (int k = 1; k < nz - 1; k++)                                                \
            for (int j = 1; j < ny - 1; j++)                                            \
                for (int i = 1; i < nx - 1; i++) {                                      \
                    const int64_t p = k * sz + j * sy + i;                              \
                    T faces = in[p - sz] + in[p + sz] + in[p - sy] + in[p + sy] +       \
                              in[p + 1] + in[p - 1];                                    \
                    T edges = -in[p - sz - sy] - in[p + sz - sy] - in[p - sz + sy] -    \
                              in[p + sz + sy] + in[p - sz - 1] + in[p + sz - 1] +       \
                              in[p - sy - 1] + in[p + sy - 1] + in[p - sz + 1] +        \
                              in[p + sz + 1] + in[p - sy + 1] + in[p + sy + 1];         \
                    out[p] = (T)(2.666 * (double)in[p] - 0.166 * (double)faces +        \
                                 0.0833 * (double)edges);                               \
                }                                                                       \
    }
ORC_POISSON(float, orc_poisson_ref_f32)
ORC_POISSON(double, orc_poisson_ref_f64)

/* ------------------------------------------------------------------------- */
/* 19-point Jacobi sweep for the Laplace problem with zero Dirichlet boundary */
/* (an operator this project defines; the reference has no stable iterated   */
/* Poisson).  Discrete operator, h = 1:                                       */
/*   (A u)(c) = (24 u(c) - 2*sum_faces u - sum_edges u) / 6                   */
/* (the 19-point Laplacian, consistent to second order).  Jacobi:            */
/*   u'(c) = (2*sum_faces u + sum_edges u + 6 f(c)) / 24                      */
/* i.e. faces weighted 1/12, edges 1/24 (a convex combination).               */
/* Residual of the input iterate r = f - A u, returned as sum r^2 over the   */
/* interior in FP64 (caller takes the square root).  f may be NULL (= 0).     */
/* Face order x-1,x+1,y-1,y+1,z-1,z+1; edges (y,z)(x,z)(x,y) pairs.          */
/* Output boundary (1 cell) is written as 0.                                  */
/* ------------------------------------------------------------------------- */
#define ORC_JACOBI19(T, NAME)                                                            \
    ORC_API double NAME(const T* in, T* out, const T* f, int nx, int ny, int nz) {      \
        const int64_t sy = nx, sz = (int64_t)nx * ny;                                   \
        double res2 = 0.0;                                                              \
        _Pragma("omp parallel for schedule(static) reduction(+ : res2)")               \
        for (int k = 0; k < nz; k++)                                                    \
            for (int j = 0; j < ny; j++)                                                \
                for (int i = 0; i < nx; i++) {                                          \
                    const int64_t p = k * sz + j * sy + i;                              \
                    if (k == 0 || j == 0 || i == 0 || k == nz - 1 || j == ny - 1 ||     \
                        i == nx - 1) {                                                  \
                        out[p] = 0;                                                     \
                        continue;                                                       \
                    }                                                                   \
                    double F = (double)in[p - 1] + in[p + 1] + in[p - sy] + in[p + sy] + \
                               in[p - sz] + in[p + sz];                                 \
                    double E = (double)in[p - sy - sz] + in[p + sy - sz] +              \
                               in[p - sy + sz] + in[p + sy + sz] + in[p - 1 - sz] +     \
                               in[p + 1 - sz] + in[p - 1 + sz] + in[p + 1 + sz] +       \
                               in[p - 1 - sy] + in[p + 1 - sy] + in[p - 1 + sy] +       \
                               in[p + 1 + sy];                                          \
                    double fc = f ? (double)f[p] : 0.0;                                 \
                    double r = fc - (24.0 * (double)in[p] - 2.0 * F - E) / 6.0;         \
                    res2 += r * r;                                                      \
                    out[p] = (T)((2.0 * F + E + 6.0 * fc) / 24.0);                      \
                }                                                                       \
        return res2;                                                                    \
    }
ORC_JACOBI19(float, orc_jacobi19_f32)
ORC_JACOBI19(double, orc_jacobi19_f64)

/* ------------------------------------------------------------------------- */
/* Error metrics.  Reference checkers: 2D sum|gold-gpu| over the strict       */
/* interior (stencil2D_5pt-v1.cu:221-228), 3D RMS + max over [lb,ub)^3        */
/* (common.hpp:84-127).  This project additionally uses a cancellation-robust */
/* per-point scale s(p) = sum_i |w_i x_i| (computed by the caller as the gold */
/* of |w| over |x|).  out[0]=max|d|, out[1]=rms|d|, out[2]=max |d|/s,         */
/* out[3]=sum|d|, out[4]=count of exact matches, out[5]=count.                */
/* ------------------------------------------------------------------------- */
ORC_API void orc_err_f64(const double* got, const double* ref, const double* scale, int64_t n,
                         double* out) {
    double mx = 0, s2 = 0, mrel = 0, s1 = 0, match = 0;
#pragma omp parallel for reduction(max : mx, mrel) reduction(+ : s2, s1, match)
    for (int64_t i = 0; i < n; i++) {
        double d = fabs(got[i] - ref[i]);
        if (d != d) d = INFINITY; /* an unwritten (NaN) or NaN result is never a match */
        if (d > mx) mx = d;
        s2 += d * d;
        s1 += d;
        if (d == 0) match += 1;
        double sc = scale ? scale[i] : fabs(ref[i]);
        double rel = d == 0 ? 0 : (sc > 0 ? d / sc : INFINITY);
        if (rel > mrel) mrel = rel;
    }
    out[0] = mx;
    out[1] = n ? sqrt(s2 / (double)n) : 0;
    out[2] = mrel;
    out[3] = s1;
    out[4] = match;
    out[5] = (double)n;
}

/* threads the OpenMP loops above use (the CPU-baseline core count) */
ORC_API int orc_max_threads(void) { return omp_get_max_threads(); }
