// stencil3d_bf16.hip -- 3D kernel instances for BF16 storage: BF16 -> BF16 (a
// single launch), BF16 -> FP32 and FP32 -> BF16 (the ends of a chain whose
// running state stays FP32 between launches).
#define SSAM_3D_PAIR_UNIT
#include "stencil3d_launch.hpp"
SSAM_INSTANTIATE_3D_PAIR(ssam::bf16_t, ssam::bf16_t)
SSAM_INSTANTIATE_3D_PAIR(ssam::bf16_t, float)
SSAM_INSTANTIATE_3D_PAIR(float, ssam::bf16_t)
