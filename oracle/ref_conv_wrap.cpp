// oracle/ref_conv_wrap.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Instantiates the reference's own CPU gold `Convolution<T>` from the header
// where it lies (ssai-2dconv/common.h:186-216, included by path from the
// reference tree through -I) and exports it with C linkage, so the oracle's
// restatement (oracle/gold.c) can be checked against the reference code itself.
// Nothing of the reference is copied here: this file only names the template.
// Output: oracle/_ref/libref_gold.so (git-ignored; built by oracle/Makefile only
// where the reference tree is present).
#include <cstring>
#include <cmath>
#include "common.h"

extern "C" {
__attribute__((visibility("default"))) void ref_conv2d_f32(const float* src, float* dst, int w,
                                                           int h, const float* f, int fw, int fh) {
    Convolution<float>(src, dst, w, h, f, fw, fh);
}
__attribute__((visibility("default"))) void ref_conv2d_f64(const double* src, double* dst, int w,
                                                           int h, const double* f, int fw, int fh) {
    Convolution<double>(src, dst, w, h, f, fw, fh);
}
}
