"""Differentiable polyenergetic projection in PyTorch.

`project(rho, proj)` maps material densities rho [n_mat][nz][ny][nx] (float32,
on the GPU) to the detector intensities (I, IAbsorb) of the projector's scan
and differentiates through the hand-written adjoint of libctfp (a
torch::autograd::Function in the `_ctfp_torch` extension): the gradient
is the backprojection of the polyenergetic residual
  dl/dL_m = (dl/dIA - dl/dI) * sum_E A[E][m] W_E exp(-sum_m' A[E][m'] L_m')
into per-material densities, with every ray's line integrals recomputed in
the backward pass rather than stored.

The reference has no gradient; the forward is its projection call
(CBCTSinMatPolyForwardProjGrid::computePolyForwProj,
Source/CBCTSinMatPolyForwardProjGrid.cpp:82-105).

The module needs the extension built (`make -C ct_forward_projection_amd/csrc
torch`, done by the package build); there is no fallback.
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from .projector import Projector

_ext = None


def _extension():
    global _ext
    if _ext is None:
        import torch  # noqa: F401  (loads libtorch before the extension)
        try:
            from . import _ctfp_torch
        except ImportError as e:
            raise ImportError("the _ctfp_torch extension is not built: run "
                              "`make -C ct_forward_projection_amd/csrc torch` "
                              "or __graft_entry__.build()") from e
        _ext = _ctfp_torch
    return _ext


def project(rho, proj: Projector, view_begin: int = 0, view_count: int | None = None,
            storage_dtype: int = L.CTFP_F32):
    """(I, IAbsorb) [views][nv][nu] of densities rho, differentiable in rho.

    The projector supplies the scan and the spectral tables (set_spectral);
    its resident volume is overwritten by rho on every call.  Calls that
    share one projector must not overlap on different streams."""
    ext = _extension()
    vc = -1 if view_count is None else int(view_count)
    I, IA = ext.project(rho, proj.handle, int(view_begin), vc, int(storage_dtype))
    return I, IA


def log_projection(I, I0, eps: float = 1e-12):
    """p = ln(I0 / I): the reference's log projection before its floor to
    whole photons (convertIandIAbsorbDt, CBCTPolyProjGPU.cu), kept continuous
    so that it can be differentiated."""
    import torch
    return torch.log(I0) - torch.log(torch.clamp(I, min=eps))


def wls_loss(p, p_target, weights):
    """Weighted least squares 0.5 * sum(w * (p - p_target)^2)."""
    return 0.5 * (weights * (p - p_target) ** 2).sum()


def poisson_target(I, seed: int = 0):
    """Counts ~ Poisson(I) (the measured target of a noisy scan), as float32
    tensors on I's device, with zero counts raised to one photon so that the
    log projection stays finite.  Drawn on the host with a seeded numpy
    generator so the target does not depend on the device."""
    import torch
    rng = np.random.default_rng(seed)
    lam = I.detach().double().cpu().numpy()
    counts = np.maximum(rng.poisson(lam), 1).astype(np.float32)
    return torch.from_numpy(counts).to(I.device)
