// stencil2d_lds_even_f32.hip -- LDS-tap kernel instances for f32 storage: even square filters (2r x 2r taps), radius 2..14.
#define SSAM_2D_LDS_UNIT
#include "stencil2d_dispatch.hpp"
SSAM_INSTANTIATE_2D_LDS(float, ssam::kBoxEven)
