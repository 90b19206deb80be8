// stencil2d_box_f32.hip -- 2D SSAM kernel instances for f32 storage: full (2r+1)^2 boxes and dense filters embedded in them, radius 1..14.
#define SSAM_2D_SHAPE_UNIT
#include "stencil2d_dispatch.hpp"
SSAM_INSTANTIATE_2D_SHAPE(float, ssam::kBox)
