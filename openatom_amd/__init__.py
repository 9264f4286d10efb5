"""MI355X-native orthonormalization path of OpenAtom (PairCalculator overlap,
Ortho S->T, backward T*psi / Lambda*psi) over hand-written gfx950 kernels.

The compute lives in liboa_ortho.so (C ABI: include/oa_ortho.h).  This
package is its host-side mirror: `Handle` (raw entry points), and
`PairCalculator` / `Ortho` named after the reference's chares, `Batch` (many
independent instances in grouped launches).
"""
from ._lib import (OaError, OrthoNotConverged, LIB_PATH, prec_code, lib)  # noqa: F401

__all__ = ["OaError", "OrthoNotConverged", "LIB_PATH", "prec_code", "lib", "Handle",
           "PairCalculator", "Ortho", "default_handle", "Batch", "instance_footprint"]


def __getattr__(name):
    if name in ("Handle", "PairCalculator", "Ortho", "default_handle", "Batch",
                "instance_footprint"):
        from . import ortho
        return getattr(ortho, name)
    raise AttributeError(name)
