"""The torch::autograd::Function over the adjoint (`_ctfp_torch`): forward
identical to the projector's, gradients identical to ctfp_backward (from the
kept line integrals and from a re-march, bit for bit equal), a
weighted least-squares loss on log projections checked by central
differences along random directions (every voxel at once), and a backward
that runs after the context has been handed other densities."""
import numpy as np
import pytest

import ct_forward_projection_amd as P
from ct_forward_projection_amd import autograd as A
from ct_forward_projection_amd import synthetic as S

pytestmark = pytest.mark.gpu


def rel_err(got, ref, floor_frac=1e-3):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    floor = floor_frac * np.abs(ref).max()
    return float((np.abs(got - ref) / np.maximum(np.abs(ref), floor)).max())


def small_case(n=32, views=4, det=(56, 44), seed=5):
    w, b, i = S.three_material_phantom(n, seed=seed)
    s = S.config2()
    s.n_vox, s.length_mm, s.det, s.d_size, s.views = (n, n, n), (n * 0.5,) * 3, det, 1.0, views
    return s, np.stack([w, b, i])


def test_autograd_forward_and_gradient_match_c_abi(require_gpu):
    import torch
    s, rho_np = small_case()
    kw = P.spectral_inputs(s, *s.det)
    dev = torch.device("cuda")
    rng = np.random.default_rng(1)
    c = torch.from_numpy(rng.normal(size=(s.views, s.det[1], s.det[0])).astype(np.float32)).to(dev)
    d = torch.from_numpy(rng.normal(size=c.shape).astype(np.float32)).to(dev)
    with P.Projector(P.scan_params(**s.scan_kwargs())) as proj:
        proj.set_spectral(**kw)
        rho = torch.from_numpy(rho_np).to(dev).requires_grad_(True)
        I, IA = A.project(rho, proj)
        (I * c + IA * d).sum().backward()
        proj.set_volumes(list(rho_np))
        I2, IA2 = proj.forward()
        g2 = proj.backward(c, d)
        torch.cuda.synchronize()
        assert torch.equal(I.detach(), I2) and torch.equal(IA.detach(), IA2)
        assert rel_err(rho.grad.cpu().numpy(), g2.cpu().numpy()) < 1e-4
        # one output only: the other's upstream gradient is zero
        rho.grad = None
        I, _ = A.project(rho, proj, view_begin=1, view_count=2)
        (I * c[1:3]).sum().backward()
        proj.set_volumes(list(rho_np))
        g3 = proj.backward(c[1:3].contiguous(), None, view_begin=1, view_count=2)
        assert rel_err(rho.grad.cpu().numpy(), g3.cpu().numpy()) < 1e-4
        # the backward from the kept line integrals and the one that marches
        # the volume again read the same L: equal bit for bit
        grads = []
        for keep in (True, False):
            rho.grad = None
            I, IA = A.project(rho, proj, keep_line_integrals=keep)
            (I * c + IA * d).sum().backward()
            grads.append(rho.grad.clone())
        torch.cuda.synchronize()
        assert torch.equal(grads[0], grads[1])


def test_backward_after_other_densities(require_gpu):
    """Backward lays out its own saved densities again, so a forward of other
    densities in between does not change the gradient."""
    import torch
    s, rho_np = small_case()
    kw = P.spectral_inputs(s, *s.det)
    dev = torch.device("cuda")
    with P.Projector(P.scan_params(**s.scan_kwargs())) as proj:
        proj.set_spectral(**kw)
        rho = torch.from_numpy(rho_np).to(dev).requires_grad_(True)
        I, _ = A.project(rho, proj)
        I.sum().backward()
        ref = rho.grad.clone()
        rho.grad = None
        I, _ = A.project(rho, proj)
        other = torch.from_numpy(rho_np[:, ::-1].copy()).to(dev)
        A.project(other * 1.5, proj)
        I.sum().backward()
        assert rel_err(rho.grad.cpu().numpy(), ref.cpu().numpy()) < 1e-4
        # volumes of another material count uploaded in between: backward
        # lays its own out again
        for keep in (True, False):
            rho.grad = None
            I, _ = A.project(rho, proj, keep_line_integrals=keep)
            proj.set_volumes(list(other[:2] * 1.5))
            I.sum().backward()
            assert rel_err(rho.grad.cpu().numpy(), ref.cpu().numpy()) < 1e-4, keep


def test_wls_directional_finite_differences(require_gpu):
    """l(rho) = WLS between log projections and a Poisson-noisy target (1e5
    photons per pixel); <grad l, delta> against Richardson-extrapolated
    central differences along three random directions, to 1e-3."""
    import torch
    s, rho_np = small_case(n=40, views=6, det=(72, 56))
    kw = P.spectral_inputs(s, *s.det)
    dev = torch.device("cuda")
    rng = np.random.default_rng(11)
    with P.Projector(P.scan_params(**s.scan_kwargs())) as proj:
        proj.set_spectral(**kw)
        I0 = proj.flat_field()
        truth = torch.from_numpy(rho_np).to(dev)
        It, _ = A.project(truth, proj)
        # the loss is evaluated in FP64 so its own rounding stays far below
        # the differences taken of it
        counts = A.poisson_target(It, seed=3).double()
        I0 = I0.double()
        p_t = A.log_projection(counts, I0)
        # start from a perturbed estimate so the residual is not just noise
        est = (truth * torch.from_numpy(rng.uniform(0.8, 1.2, size=rho_np.shape).astype(np.float32))
               .to(dev)).requires_grad_(True)

        def loss(r):
            I, _ = A.project(r, proj)
            return A.wls_loss(A.log_projection(I.double(), I0), p_t, counts)

        l0 = loss(est)
        l0.backward()
        g = est.grad.detach().double()
        atten = np.asarray(kw["atten"], np.float64)[np.asarray(kw["spectrum"]) > 0].max(axis=0)
        for trial in range(3):
            delta = torch.from_numpy(rng.normal(size=rho_np.shape)).to(dev)
            # keep every voxel's tau change small: scale per material by its attenuation
            delta *= torch.from_numpy(0.02 / atten).to(dev)[:, None, None, None]
            d32 = delta.float()
            with torch.no_grad():
                def q(h):
                    lp = loss(est.detach() + h * d32).item()
                    lm = loss(est.detach() - h * d32).item()
                    return (lp - lm) / (2 * h)
                fd = (4 * q(0.5) - q(1.0)) / 3
            an = float((g * d32.double()).sum())
            assert abs(fd - an) <= 1e-3 * abs(fd), (trial, fd, an)


def test_projector_closed_or_dropped_before_backward(require_gpu):
    """A graph keeps its projector: backward after close(), and after the
    last Python reference to the projector is gone, gives the gradient of
    an untouched projector; the context is released with the graph."""
    import gc
    import torch
    s, rho_np = small_case()
    kw = P.spectral_inputs(s, *s.det)
    dev = torch.device("cuda")
    with P.Projector(P.scan_params(**s.scan_kwargs())) as ref_proj:
        ref_proj.set_spectral(**kw)
        rho = torch.from_numpy(rho_np).to(dev).requires_grad_(True)
        I, _ = A.project(rho, ref_proj)
        I.sum().backward()
        ref = rho.grad.clone()
    # closed before backward
    proj = P.Projector(P.scan_params(**s.scan_kwargs()))
    proj.set_spectral(**kw)
    rho.grad = None
    I, IA = A.project(rho, proj)
    proj.close()
    with pytest.raises(P.CTFPError):
        A.project(rho, proj)
    I.sum().backward()
    torch.cuda.synchronize()
    assert torch.equal(rho.grad, ref)
    assert proj._ctx  # still held by the graph
    del I, IA  # the last references to the graph
    gc.collect()
    assert not proj._ctx
    # dropped before backward
    rho.grad = None

    def make():
        p = P.Projector(P.scan_params(**s.scan_kwargs()))
        p.set_spectral(**kw)
        return p

    I, _ = A.project(rho, make())
    gc.collect()
    I.sum().backward()
    torch.cuda.synchronize()
    assert torch.equal(rho.grad, ref)


def test_batched_module_directional_finite_differences(require_gpu):
    """PolyProjection over two phantoms with their own projectors (different
    detectors and view counts) and one shared projector: each output equals
    its own projection, and the gradient of a WLS loss summed over the batch
    agrees with central differences along a random direction to 1e-3."""
    import torch
    dev = torch.device("cuda")
    cases = [small_case(n=32, views=4, det=(56, 44), seed=5), small_case(n=32, views=5, det=(48, 40), seed=6)]
    projs, targets = [], []
    for s, rho_np in cases:
        pr = P.Projector(P.scan_params(**s.scan_kwargs()))
        pr.set_spectral(**P.spectral_inputs(s, *s.det))
        projs.append(pr)
        I0 = pr.flat_field().double()
        with torch.no_grad():
            It, _ = A.project(torch.from_numpy(rho_np).to(dev), pr)
        counts = A.poisson_target(It, seed=4).double()
        targets.append((I0, A.log_projection(counts, I0), counts))
    module = A.PolyProjection(projs)
    truth = [torch.from_numpy(r).to(dev) for _, r in cases]
    rng = np.random.default_rng(21)
    est = [(t * torch.from_numpy(rng.uniform(0.85, 1.15, size=t.shape).astype(np.float32)).to(dev))
           .requires_grad_(True) for t in truth]
    I, IA = module(est)
    assert isinstance(I, list) and len(I) == 2
    for b in range(2):
        ref_I, ref_IA = A.project(est[b].detach(), projs[b])
        assert torch.equal(I[b], ref_I) and torch.equal(IA[b], ref_IA)

    def loss(rs):
        Ib, _ = module(rs)
        return sum(A.wls_loss(A.log_projection(x.double(), t[0]), t[1], t[2]) for x, t in zip(Ib, targets))

    loss(est).backward()
    kw = P.spectral_inputs(cases[0][0], *cases[0][0].det)
    atten = np.asarray(kw["atten"], np.float64)[np.asarray(kw["spectrum"]) > 0].max(axis=0)
    deltas = []
    for t in truth:
        d = torch.from_numpy(rng.normal(size=t.shape)).to(dev) * torch.from_numpy(0.02 / atten).to(dev)[:, None, None, None]
        deltas.append(d.float())
    an = sum(float((e.grad.double() * d.double()).sum()) for e, d in zip(est, deltas))
    with torch.no_grad():
        def q(h):
            lp = loss([e.detach() + h * d for e, d in zip(est, deltas)]).item()
            lm = loss([e.detach() - h * d for e, d in zip(est, deltas)]).item()
            return (lp - lm) / (2 * h)
        fd = (4 * q(0.5) - q(1.0)) / 3
    assert abs(fd - an) <= 1e-3 * abs(fd), (fd, an)
    # one projector shared by a stacked batch of equal shapes
    shared = A.PolyProjection(projs[0])
    rb = torch.stack([truth[0], 0.9 * truth[0]])
    Is, IAs = shared(rb)
    assert Is.shape == (2, 4, 44, 56)
    r0, _ = A.project(rb[1], projs[0])
    assert torch.equal(Is[1], r0)
    for pr in projs:
        pr.close()
