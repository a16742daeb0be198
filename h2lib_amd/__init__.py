"""h2lib_amd -- MI355X-native near-field BEM quadrature and H2-matrix
products, behind the reference (H2Lib) operator interface.

Compute runs in libh2hip.so (hand-written HIP for gfx950).  This package is
the host-side mirror of the reference's interface for that path.
"""
from ._lib import (H2HipError, get_reproducible, hip,  # noqa: F401
                   set_reproducible)
from .device import DeviceArray, Event, Stream  # noqa: F401
from .bem3d import (BASIS_CONSTANT_BEM3D, BASIS_LINEAR_BEM3D,  # noqa: F401
                    Bem3d, NearfieldPlan, Surface3d, new_adlp_helmholtz_bem3d,
                    new_adlp_laplace_bem3d, new_dlp_helmholtz_bem3d,
                    new_dlp_laplace_bem3d, new_slp_helmholtz_bem3d,
                    new_slp_laplace_bem3d)
