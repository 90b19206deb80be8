// stencil3d_f32f64.hip -- 3D kernel instances for F32 and F64 storage (one and
// two sweeps per launch); F32 -> F32 also carries the middle launches of a
// BF16 / FP16 chain's FP32 running state.
#define SSAM_3D_PAIR_UNIT
#include "stencil3d_launch.hpp"
SSAM_INSTANTIATE_3D_PAIR(float, float)
SSAM_INSTANTIATE_3D_PAIR(double, double)
