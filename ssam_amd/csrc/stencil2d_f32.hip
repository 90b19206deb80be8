// stencil2d_f32.hip -- 2D SSAM kernel instances for f32 storage.
#include "stencil2d_dispatch.hpp"
SSAM_INSTANTIATE_2D(float)
