// stencil2d_bf16.hip -- 2D SSAM kernel instances for bf16 storage.
#include "stencil2d_dispatch.hpp"
SSAM_INSTANTIATE_2D(ssam::bf16_t)
