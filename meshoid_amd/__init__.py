"""meshoid_amd -- MI355X (gfx950) implementation of meshoid's neighbour-weighted
operator path: adaptive smoothing lengths and densities from an exact kNN, and
least-squares derivative operators (D, D2, Div, Curl), behind the reference's
``Meshoid`` API.  All computation runs in libmeshoid_hip.so (hand-written HIP);
importing this package fails if that library is missing.
"""
from . import engine
from ._lib import lib as _libhandle
from .functions import (GridDensity, GridSurfaceDensity, HsmlIter, Kernel, gradient_weights, nearest_image,
                        nearest_image_v)
from .meshoid import Meshoid
from .radiation import radtransfer

__version__ = "1.0.0"
__all__ = ["Meshoid", "HsmlIter", "Kernel", "gradient_weights", "nearest_image",
           "nearest_image_v", "GridSurfaceDensity", "GridDensity", "radtransfer", "engine"]
