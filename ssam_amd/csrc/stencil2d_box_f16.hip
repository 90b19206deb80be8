// stencil2d_box_f16.hip -- 2D SSAM kernel instances for f16 storage: full (2r+1)^2 boxes and dense filters embedded in them, radius 1..14.
#define SSAM_2D_SHAPE_UNIT
#include "stencil2d_dispatch.hpp"
SSAM_INSTANTIATE_2D_SHAPE(ssam::f16_t, ssam::kBox)
