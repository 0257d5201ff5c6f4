"""cs_dmri_amd — MI355X (gfx950) 3D-SHORE fit with the interface of PennLINC/cs_dmri.

The compute path is ``libshorefit.so`` (hand-written HIP kernels behind the C
ABI in include/shorefit.h); there is no CPU fallback.
"""

from .models import (design_matrix, fista_batch, graddev_gram, graddev_tune, graddev_variants,  # noqa: F401
                     n_coeffs, release_device_cache, shore_fit, shore_microstructure, shore_odf_sh, shore_predict,
                     tikhonov_batch, tournier_matrix)
from .shore import GradientTable, ShoreFit, ShoreModel  # noqa: F401

__version__ = "0.1.0"
