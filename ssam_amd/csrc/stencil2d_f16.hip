// stencil2d_f16.hip -- 2D SSAM kernel instances for f16 storage.
#include "stencil2d_dispatch.hpp"
SSAM_INSTANTIATE_2D(ssam::f16_t)
