// ssai-3dstencil <7pt|13pt|poisson> <float|double> [N]
//
// Keeps the command line and output of the reference 3D driver
// (ssai-3dstencil/main.cpp:11-31) and its harnesses (tmp3d7pt-v2.cu `_3d7pt`,
// tmp3d13pt.cu `_3d13pt`, tmpPoisson.cu `_poisson`): an N^3 grid (default 512,
// GRID_SIZE of common.hpp:9-10) of U[0,1) + 0.02121 values, the reference
// coefficients, the reference's iteration counts (the non-V100 tables,
// tmp3d7pt-v2.cu:369-388, tmp3d13pt.cu:405-409, tmpPoisson.cu:498-512), the timing
// line with GUPS, and the RMS check over [r*IT, N - r*IT)^3 against the CPU
// iteration (common.hpp:84-127, TOLERANCE 1e-5).
#include <cstdio>
#include <cstring>
#include <vector>

#include "harness.hpp"

using namespace harness;

namespace {

// descriptor tap order (include/ssam.h) with the reference coefficients
const double k7[7] = {0.3, 0.5, 0.6, 0.2, 0.4, 0.1, 0.7};
const double k13[13] = {0.3, 1.1, 1.2, 0.6, 0.8, 0.2, 0.4, 1.0, 1.3, 0.7, 0.9, 0.1, 0.5};
const double kp[19] = {2.666,   -0.166,  -0.166,  -0.166,  -0.166, -0.166, -0.166,
                       -0.0833, -0.0833, -0.0833, -0.0833, 0.0833, 0.0833, 0.0833,
                       0.0833,  0.0833,  0.0833,  0.0833,  0.0833};

struct Op {
    int shape, radius, ntaps;
    const double* c;
    int flop;
};

// (dz, dy, dx) of tap t in the descriptor order
void tap(const Op& op, int t, int& dz, int& dy, int& dx) {
    dz = dy = dx = 0;
    if (op.shape == SSAM_SHAPE_FACE_EDGE && t >= 7) {
        const int e = t - 7, s0 = (e & 1) ? 1 : -1, s1 = (e & 2) ? 1 : -1;
        if (e < 4) { dz = s1; dy = s0; }
        else if (e < 8) { dz = s1; dx = s0; }
        else { dy = s1; dx = s0; }
        return;
    }
    if (t == 0) return;
    const int k = (t - 1) / 6 + 1, a = ((t - 1) % 6) / 2, sg = ((t - 1) % 2) ? 1 : -1;
    (a == 0 ? dx : (a == 1 ? dy : dz)) = sg * k;
}

// CPU iteration for the self-check: interior only, boundary left at zero
template <typename T>
void sweep(const Op& op, const T* in, T* out, int n) {
    const long long sy = n, sz = static_cast<long long>(n) * n;
    std::vector<long long> off(op.ntaps);
    std::vector<T> c(op.ntaps);
    for (int t = 0; t < op.ntaps; ++t) {
        int dz, dy, dx;
        tap(op, t, dz, dy, dx);
        off[t] = dz * sz + dy * sy + dx;
        c[t] = static_cast<T>(op.c[t]);
    }
    const int r = op.radius;
#pragma omp parallel for schedule(static)
    for (int k = r; k < n - r; ++k)
        for (int j = r; j < n - r; ++j)
            for (int i = r; i < n - r; ++i) {
                const long long p = k * sz + j * sy + i;
                T acc = 0;
                for (int t = 0; t < op.ntaps; ++t) acc += c[t] * in[p + off[t]];
                out[p] = acc;
            }
}

template <typename T>
int run(const Op& op, int n, int iters, int cache) {
    const size_t cnt = static_cast<size_t>(n) * n * n;
    std::vector<T> in(cnt), out(cnt), gold(cnt, T(0)), tmp(cnt, T(0));
    for (size_t i = 0; i < cnt; ++i) in[i] = static_cast<T>(u01(1, i) + 0.02121);
    DevBuf<T> din(cnt), dout(cnt), dwork(iters > 1 ? cnt : 1);
    din.upload(in.data());
    ssam_stencil3d_desc d{};
    d.storage = sizeof(T) == 8 ? SSAM_F64 : SSAM_F32;
    d.shape = op.shape;
    d.radius = op.radius;
    d.steps = iters;
    d.boundary = SSAM_BOUNDARY_DIRICHLET0;
    d.flags = 0;
    d.coeffs = op.c;
    SSAM_OK_OR_DIE(ssam_stencil3d(&d, din.p, dout.p, n, n, n, dwork.p, nullptr));
    HIP_OK(hipDeviceSynchronize());
    EventTimer tm;
    tm.start();
    SSAM_OK_OR_DIE(ssam_stencil3d(&d, din.p, dout.p, n, n, n, dwork.p, nullptr));
    const float ftime = tm.stop_ms();
    dout.download(out.data());
    const double gups = double(n) * n * n * 0.000001 / ftime * iters;
    std::printf("dtype=%s, all_time=%fms, one_itr_time=%fms, flop = %d, N=%d,Iterative=%d, "
                "CacheCount=%d, WarpCount=%d,  gups=%.2fGUPS, gflop=%.2fGFLOP/s\n",
                sizeof(T) == 8 ? "double" : "float", ftime, ftime / iters, op.flop, n, iters,
                cache, 4, gups, gups * op.flop);

    std::vector<T> cur(in);
    for (int s = 0; s < iters; ++s) {
        std::fill(tmp.begin(), tmp.end(), T(0));
        sweep(op, cur.data(), tmp.data(), n);
        cur.swap(tmp);
    }
    const int lb = op.radius * iters, ub = n - op.radius * iters;
    double err = 0, maxe = 1e-5, sum = 0;
    long long match = 0, unmatch = 0;
    for (int k = lb; k < ub; ++k)
        for (int j = lb; j < ub; ++j)
            for (int i = lb; i < ub; ++i) {
                const size_t p = (static_cast<size_t>(k) * n + j) * n + i;
                sum += out[p];
                const double e = std::fabs(static_cast<double>(out[p]) - static_cast<double>(cur[p]));
                (e == 0 ? match : unmatch)++;
                err += e * e;
                if (e > maxe) maxe = e;
            }
    std::printf("sum = %e, match = %lld, unmatch = %lld\n", sum, match, unmatch);
    std::printf("[Test] Max Error : %e\n", maxe);
    const double vol = double(ub - lb) * (ub - lb) * (ub - lb);
    const double rms = vol > 0 ? std::sqrt(err / vol) : 0.0;
    std::printf("N=%d, [Test] RMS Error : %e\n", n, rms);
    return rms > 1e-5 ? -1 : 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::printf("error, argc = %d\n", argc);
        return 1;
    }
    const char* fn = argv[1];
    const bool dbl = std::strstr(argv[2], "double") != nullptr;
    const int n = argc > 3 ? std::atoi(argv[3]) : 512;
    Op op;
    int its[4] = {0, 0, 0, 0};
    int cache = 0;
    if (std::strstr(fn, "7pt")) {
        op = Op{SSAM_SHAPE_STAR, 1, 7, k7, 13};
        its[0] = 2; its[1] = 3; its[2] = 3; its[3] = 4;
        cache = dbl ? 14 : 18;
    } else if (std::strstr(fn, "13pt")) {
        op = Op{SSAM_SHAPE_STAR, 2, 13, k13, 25};
        its[0] = 2;
        cache = dbl ? 16 : 20;
    } else if (std::strstr(fn, "poisson")) {
        op = Op{SSAM_SHAPE_FACE_EDGE, 1, 19, kp, 21};
        its[0] = 2; its[1] = 3; its[2] = 3;
        cache = dbl ? 16 : 20;
    } else {
        std::printf("invalid function name!\n");
        return 0;
    }
    std::printf("use MI355X GPU, ");
    for (int i = 0; i < 4 && its[i]; ++i) {
        std::printf("%s:%s:Datatype:%d\n", __FILE__, fn, dbl ? 8 : 4);
        if (dbl) run<double>(op, n, its[i], cache);
        else run<float>(op, n, its[i], cache);
    }
    return 0;
}
