"""The committed fixture tests/golden/small_scan.npz (written by
tests/golden/make_golden.py): the oracle must keep reproducing it, and the HIP
path must match it to the FP32 parity bound."""
import os

import numpy as np
import pytest

from oracle import oracle as O

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "small_scan.npz")


@pytest.fixture(scope="module")
def fx():
    return dict(np.load(FIX, allow_pickle=False))


def _geom(fx):
    n, _, _, lx, ly, lz, views, nu, nv, sod, sdd, ds, rot = fx["geom"].tolist()
    return O.geom(int(n), int(n), int(n), lx, ly, lz, int(views), int(nu), int(nv), sod, sdd, ds,
                  int(rot))


def _det(fx):
    return np.stack([O.det_response(1.0, float(m), fx["thickness"]) for m in fx["scint_mu"]])


def test_oracle_reproduces_fixture(fx):
    g = _geom(fx)
    p = O.paras(g)
    L = O.march_f32(g, p, list(fx["vols"]))
    np.testing.assert_array_equal(L, fx["L"])
    I, IA, I0 = O.spectrum_f32(L, fx["atten"], fx["spectrum"], float(fx["i0_val"]), fx["grid"], _det(fx))
    np.testing.assert_array_equal(I, fx["I"])
    np.testing.assert_array_equal(I0, fx["I0"])
    Iu, IAu, proj = O.postprocess(I, IA, I0)
    np.testing.assert_array_equal(Iu, fx["Iu"])
    # FP32 mirror vs FP64 evaluation of the same method
    assert np.abs(fx["I"] - fx["I64"]).max() / fx["I64"].max() < 1e-5


@pytest.mark.gpu
def test_hip_path_matches_fixture(require_gpu, fx):
    import torch
    import ct_forward_projection_amd as P
    g = _geom(fx)
    scan = P.scan_params(p_num_x=g.nx, p_num_y=g.ny, p_num_z=g.nz, p_length_x=g.lx,
                         p_length_y=g.ly, p_length_z=g.lz, proj_num=g.n_views, d_num_u=g.nu,
                         d_num_v=g.nv, sod=g.sod, sdd=g.sdd, rotated_direction=g.rot,
                         d_size=g.dsize)
    with P.Projector(scan) as proj:
        proj.set_volumes(list(fx["vols"]))
        proj.set_spectral(fx["spectrum"], fx["atten"], float(fx["i0_val"]), grid=fx["grid"],
                          scint_mu=fx["scint_mu"], scint_thickness=fx["thickness"])
        I, IA, L = proj.forward(want_L=True)
        I0 = proj.flat_field()
        Iu, IAu, pr = P.postprocess(I, IA, I0)
        torch.cuda.synchronize()
        I, L, I0, Iu, pr = (t.cpu().numpy() for t in (I, L, I0, Iu, pr))
    # relative to the array maximum: the fixture's FP32 mirror carries the
    # reference's FP32 sample-position noise, the HIP path sets geometry in FP64
    rel = lambda a, b: np.abs(a.astype(np.float64) - b).max() / np.abs(b).max()
    assert rel(L, fx["L64"]) < 1e-5 and rel(L, fx["L"]) < 1e-4
    assert rel(I, fx["I64"]) < 1e-5 and rel(I, fx["I"]) < 1e-4
    assert rel(I0, fx["I0"]) < 1e-6
    assert np.abs(Iu.view(np.uint32).astype(np.int64) - fx["Iu"].astype(np.int64)).max() <= 1
    assert rel(pr, fx["proj"]) < 1e-3
