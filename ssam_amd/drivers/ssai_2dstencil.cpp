// ssai-2dstencil <size> <pts> <float|double>
//   pts in {5, 9, 13, 17, 21, 25, 64, 81, 121, tmp_5, tmp_9}
//
// Keeps the command line and output lines of the reference driver
// (ssai-2dstencil/main.cpp:20-56 dispatching stencil_<pts>_<dtype>) and of its
// harnesses (stencil2D_*pt.cu `Test`, stencil2d_temp_{5,9}pt.cu `Test`): input from
// ../data/Lena<N>x<N>.raw when present, else integer values 0..255 (temporal:
// U[0,1) + 0.02121); the reference filter tables; one timed launch through the C
// ABI; the CPU self-check over the strict interior.  The reference draws its
// inputs from rand(); here they come from a seeded splitmix64 stream.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "harness.hpp"
#include "ref_tables.hpp"

using namespace harness;

namespace {

struct Table {
    int k;                   // filter side
    std::vector<double> w;   // k*k row-major
};

Table make(const double* w, int k) { return Table{k, std::vector<double>(w, w + k * k)}; }

Table star_int(int r) {
    switch (r) {
        case 1: return make(ref_tables::star_r1, ref_tables::star_r1_k);
        case 2: return make(ref_tables::star_r2, ref_tables::star_r2_k);
        case 3: return make(ref_tables::star_r3, ref_tables::star_r3_k);
        case 4: return make(ref_tables::star_r4, ref_tables::star_r4_k);
        case 5: return make(ref_tables::star_r5, ref_tables::star_r5_k);
        default: return make(ref_tables::star_r6, ref_tables::star_r6_k);
    }
}

Table box_table(int pts) {
    if (pts == 64) return make(ref_tables::box64, ref_tables::box64_k);
    if (pts == 81) return make(ref_tables::box81, ref_tables::box81_k);
    return make(ref_tables::box121, ref_tables::box121_k);
}

template <typename T>
int run(int size, const std::string& pts) {
    const bool temporal = pts.rfind("tmp_", 0) == 0;
    int steps = 1, shape = SSAM_SHAPE_STAR;
    Table tb;
    if (pts == "tmp_5") {
        steps = 5;  // stencil2d_temp_5pt.cu:226 fuses R=5 steps
        tb = make(ref_tables::temporal5, ref_tables::temporal5_k);
    } else if (pts == "tmp_9") {
        steps = 3;  // stencil2d_temp_9pt.cu:235 fuses R=3 steps
        tb = make(ref_tables::temporal9, ref_tables::temporal9_k);
    } else {
        const int p = std::atoi(pts.c_str());
        if (p == 64 || p == 81 || p == 121) {
            shape = SSAM_SHAPE_BOX;
            tb = box_table(p);
        } else if (p >= 5 && p <= 25 && (p - 1) % 4 == 0) {
            tb = star_int((p - 1) / 4);
        } else {
            std::printf("error, do not call functions\n");
            return 0;
        }
    }
    const int width = size, height = size;
    Image<T> img;
    char path[1024];
    std::snprintf(path, sizeof(path), "../data/Lena%dx%d.raw", width, height);
    if (!img.load_uchar(path, width, height)) {
        std::printf("Load failed : %s, generate random data\n", path);
        img.alloc(width, height);
        for (size_t i = 0; i < img.data.size(); ++i)
            img.data[i] = temporal ? static_cast<T>(u01(0, i) + 0.02121)
                                   : static_cast<T>(std::floor(256.0 * u01(0, i)));
    } else {
        std::printf("Load success : %s\n", path);
    }

    DevBuf<T> dsrc(img.data.size()), ddst(img.data.size()), dwork(steps > 1 ? img.data.size() : 1);
    dsrc.upload(img.data.data());
    std::vector<double> coef(tb.w);
    ssam_stencil2d_desc d{};
    d.storage = sizeof(T) == 8 ? SSAM_F64 : SSAM_F32;
    d.shape = shape;
    d.fw = d.fh = tb.k;
    d.steps = steps;
    d.boundary = SSAM_BOUNDARY_ZERO;
    d.flags = 0;
    d.coeffs = coef.data();
    // warm-up launch (module load), then the timed one
    SSAM_OK_OR_DIE(ssam_stencil2d(&d, dsrc.p, ddst.p, width, height, width, dwork.p, nullptr));
    HIP_OK(hipDeviceSynchronize());
    EventTimer tm;
    tm.start();
    SSAM_OK_OR_DIE(ssam_stencil2d(&d, dsrc.p, ddst.p, width, height, width, dwork.p, nullptr));
    const float inc = tm.stop_ms();
    Image<T> out;
    out.alloc(width, height);
    ddst.download(out.data.data());

    const int R = tb.k / 2;
    if (temporal) {
        const double flops = double(width) * height * steps * 0.000001 / inc * (pts == "tmp_5" ? 9.0 : 17.0);
        std::printf("%dx%d , %dx%d , proc_count=%d, cache=%d, Iterative=%d, BLOCK_SIZE=%d, %f ms , %f fps, flops=%lf\n",
                    width, height, tb.k, tb.k, 0, 2 * R * steps + 1, steps, 256, inc, 1000.0 / inc, flops);
    } else {
        std::printf("%dx%d , %dx%d , proc_count=%d, cache=%d, BLOCK_SIZE=%d, %f ms , %f fps\n", width,
                    height, tb.k, tb.k, 0, 2 * R + 1, 256, inc, 1000.0 / inc);
    }

    // CPU self-check over the strict interior (stencil2D_5pt-v1.cu:221-228)
    std::vector<T> ref(img.data), tmp(img.data.size());
    std::vector<T> f(tb.w.begin(), tb.w.end());
    for (int s = 0; s < steps; ++s) {
        correlate(ref.data(), tmp.data(), width, height, f.data(), tb.k, tb.k);
        ref.swap(tmp);
    }
    const int m = R * steps;
    double dif = 0;
    for (int y = m + 1; y < height - m; ++y)
        for (int x = m + 1; x < width - m; ++x) {
            const size_t i = static_cast<size_t>(y) * width + x;
            dif += std::fabs(static_cast<double>(ref[i]) - static_cast<double>(out.data[i]));
        }
    if (temporal) {
        const double avg = dif / ((double(width) - tb.k + 1) * (double(height) - tb.k + 1));
        std::printf("verify dif =%e, precision=%e\n", dif, avg);
    } else {
        std::printf("verify dif =%f\n", dif);
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    for (int i = 0; i < argc; i++) {
        std::printf("%s ", argv[i]);
        if (i == argc - 1) std::printf("\n");
    }
    if (argc < 4) {
        std::printf("error, do not call functions\n");
        return 0;
    }
    const int size = std::atoi(argv[1]);
    const std::string pts = argv[2];
    const char* dtype = argv[3];
    std::printf("call_function : stencil_%spt_%s : \n", pts.c_str(), dtype);
    if (std::strcmp(dtype, "double") == 0) return run<double>(size, pts);
    if (std::strcmp(dtype, "float") == 0) return run<float>(size, pts);
    std::printf("error!\n");
    return 0;
}
