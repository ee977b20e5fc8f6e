"""The differentiable projector at its full workload: 3 x 256^3 material
densities (1 mm voxels), a 512 x 512 detector at 0.8 mm, 180 views, 120 kVp
(101 active bins), Pb grid and CsI response (the projection call this
differentiates: Source/CBCTPolyForwardProjGrid.cpp:108-203).

  - forward against the FP64 oracle on two views far apart (1e-4);
  - the gather adjoint against central differences of the GPU forward at 64
    random voxels, with the loss over all 180 views (1e-3);
  - the adjoint against the FP64 CPU adjoint (exact backprojection of the
    FP64 spectrum chain rule), with the upstream gradient nonzero on two
    views only so that the CPU side stays a few seconds (1e-4)."""
import numpy as np
import pytest

import ct_forward_projection_amd as P
from ct_forward_projection_amd import synthetic as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel_err(got, ref, floor_frac):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    return float((np.abs(got - ref) / np.maximum(np.abs(ref), floor_frac * np.abs(ref).max())).max())


@pytest.fixture(scope="module")
def case():
    s = S.config4()
    vols = [np.ascontiguousarray(v) for v in S.three_material_phantom(256, seed=41)]
    kw = P.spectral_inputs(s, *s.det)
    det = np.stack([O.det_response(kw["det_factor"], float(m), kw["scint_thickness"]) for m in kw["scint_mu"]])
    return s, vols, kw, det


def test_config4_forward_two_views(require_gpu, case):
    import torch
    s, vols, kw, det = case
    g = O.geom_from_setup(s)
    p = O.paras(g)
    assert p.n_steps == 541
    with P.Projector(P.scan_params(**s.scan_kwargs())) as proj:
        proj.set_volumes(vols)
        proj.set_spectral(**kw)
        I0 = proj.flat_field()
        I, IA, L = proj.forward(want_L=True)
        torch.cuda.synchronize()
        I, IA, L, I0 = (t.cpu().numpy() for t in (I, IA, L, I0))
    for v in (17, 123):
        print(f"config4 forward: view {v} against FP64", flush=True)
        L64 = O.march_f64(g, p, vols, v, 1)
        I64, IA64, I064 = O.spectrum_f64(L64, kw["atten"], kw["spectrum"], kw["i0_val"], kw["grid"], det)
        assert rel_err(L[v:v + 1], L64, 1e-3) < 1e-4
        assert rel_err(I[v:v + 1], I64, 1e-3) < 1e-4
        assert rel_err(IA[v:v + 1], IA64, 1e-3) < 1e-4
        pg = np.log(I0[None].astype(np.float64) / I[v:v + 1])
        assert rel_err(pg, np.log(I064[None] / I64), 1e-2) < 1e-4
    assert (L[:, 0] > 0).mean() > 0.3  # the phantom fills the field of view


def test_config4_finite_differences(require_gpu, case):
    """dl/d(rho) at 64 random voxels inside the phantom's support against
    Richardson-extrapolated central differences of the GPU forward, with
    l = sum(c * I) + sum(d * IAbsorb) over all 180 views (c, d seeded normal
    per pixel).  Steps and metric as in the small-scan test."""
    import torch
    s, vols, kw, _ = case
    dev = torch.device("cuda")
    gen = torch.Generator(device=dev).manual_seed(7)
    shape = (s.views, s.det[1], s.det[0])
    c = torch.randn(shape, generator=gen, device=dev, dtype=torch.float64)
    d = torch.randn(shape, generator=gen, device=dev, dtype=torch.float64)
    vol_t = torch.from_numpy(np.stack(vols)).to(dev)
    with P.Projector(P.scan_params(**s.scan_kwargs())) as proj:
        proj.set_spectral(**kw)
        proj.set_volumes(list(vol_t))
        grad = proj.backward(c.float(), d.float())
        torch.cuda.synchronize()
        grad = grad.cpu().numpy().astype(np.float64)

        def loss():
            proj.set_volumes(list(vol_t))
            I, IA = proj.forward()
            return float((I.double() * c).sum() + (IA.double() * d).sum())

        def quotient(m, z, y, x, h):
            old = vol_t[m, z, y, x].item()
            vol_t[m, z, y, x] = old + h
            hp = vol_t[m, z, y, x].item()
            lp = loss()
            vol_t[m, z, y, x] = old - h
            hm = vol_t[m, z, y, x].item()
            lm = loss()
            vol_t[m, z, y, x] = old
            return (lp - lm) / (hp - hm)

        active = np.asarray(kw["spectrum"]) > 0
        step = 0.4 / np.asarray(kw["atten"], np.float64)[active].max(axis=0)
        rng = np.random.default_rng(19)
        errs = []
        while len(errs) < 64:
            m = int(rng.integers(0, 3))
            z, y, x = (int(v) for v in rng.integers(8, 248, size=3))
            if grad[m, z, y, x] == 0.0:
                continue
            h = float(np.float32(step[m]))
            fd = (4 * quotient(m, z, y, x, h / 2) - quotient(m, z, y, x, h)) / 3
            errs.append(abs(fd - grad[m, z, y, x]) / max(abs(fd), 1e-3 * np.abs(grad[m]).max()))
            if len(errs) % 16 == 0:
                print(f"config4 finite differences: {len(errs)} of 64 voxels, max err {max(errs):.2e}",
                      flush=True)
    assert max(errs) < 1e-3, (max(errs), np.median(errs))


def test_config4_adjoint_vs_fp64_two_views(require_gpu, case):
    import torch
    s, vols, kw, det = case
    views = (40, 130)
    rng = np.random.default_rng(23)
    shape = (s.views, s.det[1], s.det[0])
    gI = np.zeros(shape, np.float32)
    gIA = np.zeros(shape, np.float32)
    for v in views:
        gI[v] = rng.normal(size=shape[1:])
        gIA[v] = rng.normal(size=shape[1:])
    dev = torch.device("cuda")
    with P.Projector(P.scan_params(**s.scan_kwargs())) as proj:
        proj.set_volumes(vols)
        proj.set_spectral(**kw)
        grad = proj.backward(torch.from_numpy(gI).to(dev), torch.from_numpy(gIA).to(dev))
        torch.cuda.synchronize()
        grad = grad.cpu().numpy()
    g = O.geom_from_setup(s)
    p = O.paras(g)
    ref = np.zeros((3,) + vols[0].shape)
    scale = np.zeros_like(ref)
    for v in views:
        print(f"config4 adjoint: FP64 backprojection of view {v}", flush=True)
        L64 = O.march_f64(g, p, vols, v, 1)
        gL = O.spectrum_grad_f64(L64, kw["atten"], kw["spectrum"], kw["i0_val"], gI[v:v + 1], gIA[v:v + 1],
                                 kw["grid"], det)
        # the gradient and the backprojection of |dl/dL| (the scale of each
        # voxel's cancelling sum) in one pass, as six channels
        both = O.backproject_f64(g, p, np.concatenate([gL, np.abs(gL)], axis=1), view0=v)
        ref += both[:3]
        scale += both[3:]
    err = np.abs(grad - ref) / np.maximum(scale, 1e-3 * scale.max())
    assert err.max() < 1e-4, err.max()
