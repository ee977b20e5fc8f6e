"""Scans of more than four materials (the reference reads any MaterialNum,
Source/ForwardProjection.cpp:290-316; Source/PhantomObject.cpp:38-58): the
channels are marched in groups of four and joined by a spectrum pass.  The
forward, the line integrals, the adjoint and batches against the FP64 oracle;
the matrix-core spectrum with attenuation up to the 1-keV bins of a water
table; NaN / infinity inputs rejected."""
import numpy as np
import pytest

import ct_forward_projection_amd as P
from ct_forward_projection_amd import _lib as L
from ct_forward_projection_amd import synthetic as S
from oracle import oracle as O

from test_gpu_parity import MIXED_TOL, _errors, check_all, log_errors, oracle_expect, to_np

pytestmark = pytest.mark.gpu


def six_material_case(n=40, views=6, det=(72, 56), d_size=1.0):
    w, b, i = S.three_material_phantom(n, seed=61)
    w2, b2, i2 = S.three_material_phantom(n, seed=62)
    vols = [w, b, i, 0.3 * w2, 0.5 * b2, 2.0 * i2]
    s = S.config2()
    s.n_vox, s.length_mm, s.det, s.d_size, s.views = (n, n, n), (n * 0.5,) * 3, det, d_size, views
    s.materials = (S.WATER, S.BONE, S.IODINE, S.ALUMINIUM, S.CAESIUM, S.IODINE)
    s.scint_thickness_err = 0.05
    return s, [np.ascontiguousarray(v, dtype=np.float32) for v in vols]


def project(s, vols, kw, dtype=P.CTFP_F32, march="lds"):
    import torch
    with P.Projector(P.scan_params(**s.scan_kwargs())) as proj:
        proj.set_volumes(vols, storage_dtype=dtype)
        proj.set_spectral(**kw)
        proj.set_tuning(march=march)
        I, IA, Lf = proj.forward(want_L=True)
        I0 = proj.flat_field()
        Ll = proj.line_integrals()
        I2, IA2 = proj.forward()
        torch.cuda.synchronize()
        # with or without L requested, and L from ctfp_line_integrals: the same bits
        assert torch.equal(I, I2) and torch.equal(IA, IA2) and torch.equal(Lf, Ll)
        return to_np(I), to_np(IA), to_np(Lf), to_np(I0)


def test_six_materials_forward_fp32(require_gpu):
    s, vols = six_material_case()
    kw = P.spectral_inputs(s, *s.det)
    got = project(s, vols, kw)
    assert got[2].shape == (s.views, 6, s.det[1], s.det[0])
    check_all(got, oracle_expect(s, vols, kw, 0, s.views), name="six_materials_fp32")
    # the global-gather march gives the same bits as the LDS-staged one
    glob = project(s, vols, kw, march="global")
    for a, b in zip(got, glob):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("dtype,name", [(P.CTFP_F16, "f16"), (P.CTFP_BF16, "bf16")])
def test_six_materials_forward_16bit(require_gpu, dtype, name):
    s, vols = six_material_case()
    kw = P.spectral_inputs(s, *s.det)
    got = project(s, vols, kw, dtype=dtype)
    _, ref64 = oracle_expect(s, vols, kw, 0, s.views)
    e = _errors(got, ref64)
    log_errors(f"six_materials_{name}", vs_fp64=e)
    for k in ("L", "I", "IA", "p"):
        assert e[k] < MIXED_TOL[name], (k, e)


@pytest.mark.parametrize("mode", ["gather", "cell"])
def test_six_materials_adjoint(require_gpu, mode):
    """The voxel and cell gather adjoints of a 6-material scan (two channel
    groups) against the FP64 CPU adjoint, with L recomputed and with L given;
    the scatter adjoint (up to four materials) refuses it."""
    import torch
    s, vols = six_material_case(n=32, views=5, det=(56, 44))
    kw = P.spectral_inputs(s, *s.det)
    rng = np.random.default_rng(71)
    gI = rng.normal(size=(s.views, s.det[1], s.det[0])).astype(np.float32)
    gIA = rng.normal(size=gI.shape).astype(np.float32)
    dev = torch.device("cuda")
    with P.Projector(P.scan_params(**s.scan_kwargs())) as proj:
        proj.set_volumes(vols)
        proj.set_spectral(**kw)
        proj.set_tuning(adjoint=mode)
        _, _, Lg = proj.forward(want_L=True)
        g1 = proj.backward(torch.from_numpy(gI).to(dev), torch.from_numpy(gIA).to(dev))
        g2 = proj.backward(torch.from_numpy(gI).to(dev), torch.from_numpy(gIA).to(dev), L_in=Lg)
        torch.cuda.synchronize()
        assert torch.equal(g1, g2)
        grad = to_np(g1)
        proj.set_tuning(adjoint="scatter")
        with pytest.raises(P.CTFPError) as e:
            proj.backward(torch.from_numpy(gI).to(dev))
        assert e.value.code == L.CTFP_ERR_UNSUPPORTED
    g = O.geom_from_setup(s)
    p = O.paras(g)
    L64 = O.march_f64(g, p, vols)
    det = np.stack([O.det_response(kw["det_factor"], float(m), kw["scint_thickness"]) for m in kw["scint_mu"]])
    gL = O.spectrum_grad_f64(L64, kw["atten"], kw["spectrum"], kw["i0_val"], gI, gIA, kw["grid"], det)
    ref = O.backproject_f64(g, p, gL)
    scale = O.backproject_f64(g, p, np.abs(gL))
    err = np.abs(grad - ref) / np.maximum(scale, 1e-3 * scale.max())
    assert err.max() < 1e-4, err.max()


def test_batch_with_six_material_item(require_gpu):
    """A batch holding a 6-material scan (run after the persistent launch in
    the same graph) and a 3-material scan: every item equals its own forward."""
    import torch
    s6, v6 = six_material_case(n=40, views=7, det=(64, 48))
    w, b, i = S.three_material_phantom(48, seed=63)
    s3 = S.config2()
    s3.n_vox, s3.length_mm, s3.det, s3.d_size, s3.views = (48,) * 3, (24.0,) * 3, (80, 64), 0.9, 9
    dev = torch.device("cuda")
    projs, items, refs = [], [], []
    for s, vols in ((s6, v6), (s3, [w, b, i])):
        pr = P.Projector(P.scan_params(**s.scan_kwargs()))
        pr.set_volumes(vols, storage_dtype=P.CTFP_F16)
        pr.set_spectral(**P.spectral_inputs(s, *s.det))
        I = torch.empty((s.views, s.det[1], s.det[0]), device=dev)
        items.append((pr, 0, s.views, I, torch.empty_like(I)))
        refs.append(pr.forward())
        projs.append(pr)
    with P.Batch(items) as batch:
        for _ in range(2):
            for it in items:
                it[3].fill_(-1.0)
                it[4].fill_(-1.0)
            batch.run()
            torch.cuda.synchronize()
            for it, (rI, rIA) in zip(items, refs):
                assert torch.equal(it[3], rI) and torch.equal(it[4], rIA)
    for pr in projs:
        pr.close()


@pytest.mark.parametrize("dtype,name", [(P.CTFP_F16, "f16"), (P.CTFP_BF16, "bf16")])
def test_spectrum_from_one_kev(require_gpu, dtype, name):
    """Unfiltered 1-120 keV spectrum: attenuation up to ~1e4 /mm per g/cm^3
    (iodine at 1 keV) enters the matrix-core spectrum of 16-bit volumes, whose
    FP16 operand scale adapts to it (no overflow to inf / NaN)."""
    import torch
    w, b, i = S.three_material_phantom(40, seed=64)
    s = S.config2()
    s.n_vox, s.length_mm, s.det, s.d_size, s.views = (40,) * 3, (20.0,) * 3, (72, 56), 0.8, 4
    s.e_min, s.al_mm = 1.0, 0.0
    kw = P.spectral_inputs(s, *s.det)
    assert kw["spectrum"][0] > 0 and float(np.max(kw["atten"][kw["spectrum"] > 0])) > 4000.0
    vols = [w, b, i]
    with P.Projector(P.scan_params(**s.scan_kwargs())) as proj:
        proj.set_volumes(vols, storage_dtype=dtype)
        proj.set_spectral(**kw)
        I, IA, Lg = proj.forward(want_L=True)
        I0 = proj.flat_field()
        torch.cuda.synchronize()
        got = to_np(I), to_np(IA), to_np(Lg), to_np(I0)
    assert np.all(np.isfinite(got[0])) and np.all(np.isfinite(got[1]))
    _, ref64 = oracle_expect(s, vols, kw, 0, s.views)
    e = _errors(got, ref64)
    log_errors(f"one_kev_{name}", vs_fp64=e)
    for k in ("L", "I", "IA", "p"):
        assert e[k] < MIXED_TOL[name], (k, e)


def test_non_finite_inputs_rejected(require_gpu):
    import torch
    s, vols = six_material_case(n=24, views=2, det=(40, 32))
    s.materials = s.materials[:3]
    vols = vols[:3]
    kw = P.spectral_inputs(s, *s.det)
    with P.Projector(P.scan_params(**s.scan_kwargs())) as proj:
        for key, where in (("spectrum", 40), ("atten", (50, 1)), ("grid", (60, 3)), ("scint_mu", 70),
                           ("scint_thickness", 5)):
            bad = dict(kw)
            arr = np.array(bad[key], dtype=np.float32, copy=True)
            arr[where] = np.nan if key != "grid" else np.inf
            bad[key] = arr
            with pytest.raises(P.CTFPError) as e:
                proj.set_spectral(**bad)
            assert e.value.code == L.CTFP_ERR_INVALID and key in str(e.value)
        proj.set_spectral(**kw)
        for dtype in (P.CTFP_F32, P.CTFP_F16):
            badv = [v.copy() for v in vols]
            badv[1][3, 4, 5] = np.inf
            with pytest.raises(P.CTFPError) as e:
                proj.set_volumes(badv, storage_dtype=dtype)
            assert e.value.code == L.CTFP_ERR_INVALID
            # the context holds no volume afterwards
            with pytest.raises(P.CTFPError) as e:
                proj.forward()
            assert e.value.code == L.CTFP_ERR_STATE
            dvol = [torch.from_numpy(v).cuda() for v in badv]
            dvol[1][0, 0, 0] = float("nan")
            with pytest.raises(P.CTFPError):
                proj.set_volumes(dvol, storage_dtype=dtype)
        proj.set_volumes(vols)
        I, IA = proj.forward()
        torch.cuda.synchronize()
        assert torch.isfinite(I).all()
