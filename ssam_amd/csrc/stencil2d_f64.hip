// stencil2d_f64.hip -- 2D SSAM kernel instances for f64 storage.
#include "stencil2d_dispatch.hpp"
SSAM_INSTANTIATE_2D(double)
