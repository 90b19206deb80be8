"""ssam_amd -- MI355X-native SSAM register-cache stencil and convolution library.

The compute is libssam.so (HIP, gfx950) behind the C ABI of include/ssam.h; this
package is its host-side binding: `ssam_amd.api` mirrors the reference harness
operators on torch device tensors, `ssam_amd.build` compiles the library.
"""
from ._lib import (BF16, BOX, CLAMP, DENSE, DIRICHLET0, F16, F32, F64, FACE_EDGE, HOLD, STAR,
                   ZERO, SsamError, lib)

__all__ = ["lib", "SsamError", "F32", "F64", "BF16", "F16", "STAR", "BOX", "DENSE",
           "FACE_EDGE", "ZERO", "CLAMP", "HOLD", "DIRICHLET0"]
__version__ = "0.1.0"
