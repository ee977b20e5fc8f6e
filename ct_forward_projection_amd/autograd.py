"""Differentiable polyenergetic projection in PyTorch.

`project(rho, proj)` maps material densities rho [n_mat][nz][ny][nx] (float32,
on the GPU) to the detector intensities (I, IAbsorb) of the projector's scan
and differentiates through the hand-written adjoint of libctfp (a
torch::autograd::Function in the `_ctfp_torch` extension): the gradient
is the backprojection of the polyenergetic residual
  dl/dL_m = (dl/dIA - dl/dI) * sum_E A[E][m] W_E exp(-sum_m' A[E][m'] L_m')
into per-material densities.  The forward keeps every ray's line integrals
(n_mat floats per ray, e.g. 566 MB for 3 materials x 180 views of 512 x 512)
so that the backward runs the adjoint from them without marching the volume
again; keep_line_integrals=False trades that memory for a re-march (and a
new layout of the saved densities) in the backward.  `project_batch` and
`PolyProjection`
do the same for several phantoms, each with its own projector (scan
geometry, detector, spectrum) or one shared projector.

The graph of every call keeps its projector: the context stays alive until
the graph is released, and Projector.close() on a projector that a graph
still uses only marks it closed (the context goes with the last graph).

The reference has no gradient; the forward is its projection call
(CBCTSinMatPolyForwardProjGrid::computePolyForwProj,
Source/CBCTSinMatPolyForwardProjGrid.cpp:82-105).

The module needs the extension, which __graft_entry__.build() compiles
(`make -C ct_forward_projection_amd/csrc torch` alone also does); there is no
fallback.
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from .projector import Projector

_ext = None
_keep = None


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


def _keep_function():
    """An identity on the densities whose graph node holds the projector and
    its graph token, so that the context outlives every graph built on it."""
    global _keep
    if _keep is None:
        import torch

        class _Keep(torch.autograd.Function):
            @staticmethod
            def forward(ctx, rho, proj, token):
                ctx.proj = proj
                ctx.token = token
                return rho.view_as(rho)

            @staticmethod
            def backward(ctx, grad):
                return grad, None, None

        _keep = _Keep
    return _keep


def project(rho, proj: Projector, view_begin: int = 0, view_count: int | None = None,
            storage_dtype: int = L.CTFP_F32, keep_line_integrals: bool = True):
    """(I, IAbsorb) [views][nv][nu] of densities rho, differentiable in rho.

    The projector supplies the scan and the spectral tables (set_spectral);
    its resident volume is overwritten by rho on every call.  Calls that
    share one projector must not overlap on different streams.  With
    keep_line_integrals (and rho requiring grad) the graph holds the line
    integrals [views][n_mat][nv][nu] until it is released."""
    ext = _extension()
    vc = -1 if view_count is None else int(view_count)
    handle = proj.handle  # raises on a closed projector
    keep = bool(keep_line_integrals) and rho.requires_grad
    if rho.requires_grad:
        rho = _keep_function().apply(rho, proj, proj._graph_token())
    I, IA = ext.project(rho, handle, int(view_begin), vc, int(storage_dtype), keep)
    return I, IA


def project_batch(rhos, projs, view_begin: int = 0, view_count: int | None = None,
                  storage_dtype: int = L.CTFP_F32, keep_line_integrals: bool = True):
    """[(I_b, IA_b)] for phantoms rhos[b] ([n_mat][nz][ny][nx] each), each
    through projs[b], or all through one projector when `projs` is a single
    Projector.  The phantoms run one after the other on the current stream
    (each launch fills the GPU); every output is differentiable in its own
    phantom, and the backward runs the adjoint per phantom."""
    rhos = list(rhos)
    plist = [projs] * len(rhos) if isinstance(projs, Projector) else list(projs)
    if len(plist) != len(rhos):
        raise ValueError(f"{len(rhos)} phantoms but {len(plist)} projectors")
    return [project(r, p, view_begin, view_count, storage_dtype, keep_line_integrals)
            for r, p in zip(rhos, plist)]


def _module_base():
    import torch
    return torch.nn.Module


class PolyProjection(_module_base()):
    """torch.nn.Module over project_batch: forward(rho) with rho a tensor
    [B][n_mat][nz][ny][nx] (or a list of B tensors) returns (I, IAbsorb),
    stacked to [B][views][nv][nu] when every projector has the same detector
    and view count, else as lists.  `projs` is one Projector shared by the
    batch or one per phantom."""

    def __init__(self, projs, view_begin: int = 0, view_count: int | None = None,
                 storage_dtype: int = L.CTFP_F32, keep_line_integrals: bool = True):
        super().__init__()
        self.projs = projs
        self.view_begin, self.view_count, self.storage_dtype = view_begin, view_count, storage_dtype
        self.keep_line_integrals = keep_line_integrals

    def forward(self, rho):
        import torch
        out = project_batch(list(rho), self.projs, self.view_begin, self.view_count, self.storage_dtype,
                            self.keep_line_integrals)
        I, IA = [o[0] for o in out], [o[1] for o in out]
        if all(t.shape == I[0].shape for t in I):
            return torch.stack(I), torch.stack(IA)
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
