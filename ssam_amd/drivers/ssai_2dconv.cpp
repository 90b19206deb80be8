// ssai-2dconv <float|double> <size> <fsize>
//
// Keeps the command line and output of the reference convolution harness
// (ssai-2dconv/conv2D.cu:175-246, `Test_conv2D`): input from ../data/Lena<N>x<N>.raw
// when present, else integers 0..255; fsize x fsize integer taps in [-10, 10]
// (conv2D.cu:115-119); the kernel through the reference-shaped entry
// ssam_conv2d_* (pitched rows, device weights) launched 100 times with a
// synchronisation after each, the average time line, the full-image check against
// the CPU correlation, and the log.conv2D.csv row (conv2D.cu:165-169).
#include <cstdio>
#include <cstring>
#include <vector>

#include "harness.hpp"

using namespace harness;

template <typename T>
static int run(int size, int fsize) {
    std::printf("file_%s : func_%s : line_%d, using dtype=%s\n", __FILE__, "Test_conv2D", __LINE__,
                sizeof(T) == 8 ? "double" : "float32");
    const int width = size, height = size;
    Image<T> img;
    char path[1024];
    std::snprintf(path, sizeof(path), "../data/Lena%dx%d.raw", width, height);
    if (!img.load_uchar(path, width, height)) {
        std::printf("Load failed : %s, generate random data\n", path);
        img.alloc(width, height);
        for (size_t i = 0; i < img.data.size(); ++i)
            img.data[i] = static_cast<T>(std::floor(256.0 * u01(1, i)));
    } else {
        std::printf("Load success : %s\n", path);
    }
    std::vector<T> filt(static_cast<size_t>(fsize) * fsize);
    for (size_t i = 0; i < filt.size(); ++i)
        filt[i] = static_cast<T>(std::floor(21.0 * u01(2, i)) - 10.0);

    // pitched device rows (the reference's cudaMalloc3D pitch, cudaLib.cuh:297-415)
    const int pitch = (width + 63) / 64 * 64;
    DevBuf<T> dsrc(static_cast<size_t>(pitch) * height), ddst(static_cast<size_t>(pitch) * height),
        dw(filt.size());
    HIP_OK(hipMemcpy2D(dsrc.p, pitch * sizeof(T), img.data.data(), width * sizeof(T),
                       width * sizeof(T), height, hipMemcpyHostToDevice));
    dw.upload(filt.data());
    auto call = [&]() {
        if constexpr (sizeof(T) == 8)
            return ssam_conv2d_f64(dsrc.p, ddst.p, width, pitch, height, dw.p, fsize, fsize, nullptr);
        else
            return ssam_conv2d_f32(dsrc.p, ddst.p, width, pitch, height, dw.p, fsize, fsize, nullptr);
    };
    const int reps = 100;
    EventTimer tm;
    tm.start();
    for (int s = 0; s < reps; ++s) {
        SSAM_OK_OR_DIE(call());
        HIP_OK(hipDeviceSynchronize());
    }
    const float inc = tm.stop_ms() / reps;
    std::vector<T> out(static_cast<size_t>(width) * height);
    HIP_OK(hipMemcpy2D(out.data(), width * sizeof(T), ddst.p, pitch * sizeof(T), width * sizeof(T),
                       height, hipMemcpyDeviceToHost));
    std::printf("%dx%d , %dx%d , proc_count=%d, cache=%d, BLOCK_SIZE=%d, %f ms , %f fps\n", width,
                height, fsize, fsize, 0, fsize, 256, inc, 1000.0 / inc);
    std::vector<T> ref(out.size());
    correlate(img.data.data(), ref.data(), width, height, filt.data(), fsize, fsize);
    double dif = 0;
    for (size_t i = 0; i < ref.size(); ++i)
        dif += std::fabs(static_cast<double>(ref[i]) - static_cast<double>(out[i]));
    std::printf("verify dif =%f\n", dif);
    FILE* fp = std::fopen("log.conv2D.csv", "at");
    if (fp) {
        std::fprintf(fp, "%dx%d, %d_%d, %d, %dx%d, %f\n", width, height, 0, fsize, 256, fsize, fsize,
                     inc);
        std::fclose(fp);
    }
    return 0;
}

int main(int argc, char** argv) {
    for (int i = 0; i < argc; i++) {
        std::printf("%s ", argv[i]);
        if (i == argc - 1) std::printf("\n");
    }
    if (argc < 4) {
        std::printf("error, argc number!");
        return 0;
    }
    std::printf("call_function : Test_2D : evaluate\n");
    const int size = std::atoi(argv[2]);
    const int fsize = std::atoi(argv[3]);
    if (std::strcmp(argv[1], "float") == 0) return run<float>(size, fsize);
    return run<double>(size, fsize);
}
