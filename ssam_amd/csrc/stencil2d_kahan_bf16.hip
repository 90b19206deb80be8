// stencil2d_kahan_bf16.hip -- 2D SSAM kernel instances for bf16 storage: compensated boxes and dense filters, radius 1..9.
#define SSAM_2D_SHAPE_UNIT
#include "stencil2d_dispatch.hpp"
SSAM_INSTANTIATE_2D_SHAPE(ssam::bf16_t, ssam::kBoxKahan)
