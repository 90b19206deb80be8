// stencil3d_f16.hip -- 3D kernel instances for FP16 storage: FP16 -> FP16 (a
// single launch), FP16 -> FP32 and FP32 -> FP16 (the ends of a chain whose
// running state stays FP32 between launches).
#define SSAM_3D_PAIR_UNIT
#include "stencil3d_launch.hpp"
SSAM_INSTANTIATE_3D_PAIR(ssam::f16_t, ssam::f16_t)
SSAM_INSTANTIATE_3D_PAIR(ssam::f16_t, float)
SSAM_INSTANTIATE_3D_PAIR(float, ssam::f16_t)
