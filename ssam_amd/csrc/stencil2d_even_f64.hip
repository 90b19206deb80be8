// stencil2d_even_f64.hip -- 2D SSAM kernel instances for f64 storage: even square filters (2r x 2r taps), radius 1..14.
#define SSAM_2D_SHAPE_UNIT
#include "stencil2d_dispatch.hpp"
SSAM_INSTANTIATE_2D_SHAPE(double, ssam::kBoxEven)
