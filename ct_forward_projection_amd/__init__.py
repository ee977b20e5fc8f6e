"""MI355X-native polyenergetic cone-beam CT forward projector.

The compute path is libctfp.so (hand-written gfx950 HIP kernels behind the C
ABI in include/ctfp.h).  This package is its Python face: ctypes bindings
(``_lib``), the projector context, batches and host models (``projector``),
the differentiable projection for PyTorch (``autograd``) and the synthetic
inputs used for measurement (``synthetic``).
"""
from ._lib import CTFP_BF16, CTFP_F16, CTFP_F32, CTFPError, lib  # noqa: F401
from .projector import (Batch, Projector, compute_paras, det_response,  # noqa: F401
                        focal_spot_offsets, grid_period_index, grid_transmission, grid_unit_heights,
                        postprocess,
                        scan_params, scint_thickness, spectral_inputs)

__all__ = [
    "Batch", "Projector", "compute_paras", "det_response", "focal_spot_offsets",
    "grid_period_index", "grid_transmission", "grid_unit_heights",
    "postprocess", "scan_params", "scint_thickness", "spectral_inputs", "lib", "CTFPError",
    "CTFP_F32", "CTFP_F16", "CTFP_BF16",
]
