// stencil2d_kahan_f32.hip -- 2D SSAM kernel instances for f32 storage: compensated boxes and dense filters, radius 1..9.
#define SSAM_2D_SHAPE_UNIT
#include "stencil2d_dispatch.hpp"
SSAM_INSTANTIATE_2D_SHAPE(float, ssam::kBoxKahan)
