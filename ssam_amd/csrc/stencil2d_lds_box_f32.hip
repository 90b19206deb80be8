// stencil2d_lds_box_f32.hip -- LDS-tap kernel instances for f32 storage: full (2r+1)^2 boxes and dense filters embedded in them, radius 2..14.
#define SSAM_2D_LDS_UNIT
#include "stencil2d_dispatch.hpp"
SSAM_INSTANTIATE_2D_LDS(float, ssam::kBox)
