"""libctfp.so without a GPU: it loads, exports every symbol include/ctfp.h
declares, and its host-side models (scan scalars, grid, scintillator) agree
with the oracle.  No compute call here touches a device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import ct_forward_projection_amd as P
from ct_forward_projection_amd import _lib as L
from ct_forward_projection_amd import synthetic as S
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "ctfp.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(ctfp_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    syms = header_symbols()
    assert len(syms) >= 20
    for name in syms:
        assert hasattr(lib, name), name
    assert set(syms) == set(L.EXPORTED)
    assert b"gfx950" in lib.ctfp_version()


@pytest.mark.parametrize("setup", [S.config1(), S.config2()])
def test_compute_paras_matches_oracle(setup):
    info = P.compute_paras(P.scan_params(**setup.scan_kwargs()))
    p = O.paras(O.geom_from_setup(setup))
    for a, b in [("fov_r", "fov_r"), ("fov_h", "fov_h"), ("dx", "dx"), ("theta_step", "theta_step"),
                 ("d_half_lu", "half_u"), ("d_half_lv", "half_v"), ("p_size_x", "px"),
                 ("p_half_z", "half_z")]:
        assert info[a] == getattr(p, b), a
    assert info["int_num"] == p.n_steps


def test_compute_paras_rejects_bad_geometry():
    for bad in (dict(p_num_x=0), dict(d_size=-1.0), dict(proj_num=0), dict(rotated_direction=2),
                dict(sdd=0.0)):
        kw = S.config1().scan_kwargs()
        kw.update(bad)
        with pytest.raises(L.CTFPError) as e:
            P.compute_paras(P.scan_params(**kw))
        assert e.value.code == L.CTFP_ERR_INVALID


@pytest.mark.parametrize("terr", [0.0, 0.05, 0.013])
def test_scint_thickness_matches_oracle(terr):
    np.testing.assert_array_equal(P.scint_thickness(4096, 0.4, terr), O.scint_thickness(4096, 0.4, terr))


def test_det_response_matches_oracle():
    t = O.scint_thickness(2048, 0.6, 0.05)
    np.testing.assert_array_equal(P.det_response(0.9, 3.1, t), O.det_response(0.9, 3.1, t))


@pytest.mark.parametrize("nu,d_mm,w,sw", [(64, 0.4, 214, 36), (63, 0.4, 214, 36),
                                          (64, 0.2, 219, 31), (61, 0.125, 91, 34),
                                          (100, 0.139, 208, 42), (32, 0.8, 220, 210)])
def test_grid_matches_oracle_bitwise(nu, d_mm, w, sw):
    """All four parity variants of the reference's unit fill, against the
    literal loop restatement; pitch truncation to whole um included."""
    mus = np.array([0.0, 3.3, 57.0, 170.0], np.float32)
    mui = np.array([0.0, 0.0, 0.2, 1.0], np.float32)
    T = P.grid_transmission(nu, d_mm, w, sw, 2.14, 1500.0, mus, mui)
    d_um = int(np.float32(d_mm) * np.float32(1000))
    for e in range(4):
        ref = O.grid_transmission(nu, d_um, w, sw, 2.14, 1500.0, float(mus[e]), float(mui[e]))
        np.testing.assert_array_equal(T[e], ref)


@pytest.mark.parametrize("nu,d_mm", [(64, 0.4), (63, 0.4), (61, 0.125)])
def test_grid_unit_heights_match_oracle(nu, d_mm):
    """The table Grid::showGridH writes (Grid.cpp:76-87, 506-523): G = nu *
    (whole-um pitch) slant heights, bit-equal to the oracle's."""
    H = P.grid_unit_heights(nu, d_mm, 214, 36, 2.14, 1500.0)
    d_um = int(np.float32(d_mm) * np.float32(1000))
    assert H.size == nu * d_um
    np.testing.assert_array_equal(H, O.grid_unit_heights(nu * d_um, 2.14, 1500.0))


def test_grid_rejects_degenerate():
    with pytest.raises(L.CTFPError):
        P.grid_transmission(2, 0.1, 214, 36, 2.0, 1500.0, [1.0], [0.0])


def test_grid_period_index():
    T = P.grid_transmission(64, 0.4, 214, 36, 2.14, 1500.0, [57.0], [0.0])[0]
    idx = P.grid_period_index(T)
    assert idx[0] == 0 and idx[-1] == 63
    inner = idx[1:-1]
    assert all(T[i - 1] > T[i] < T[i + 1] for i in inner)
    expect = [i for i in range(1, 63) if T[i - 1] > T[i] and T[i] < T[i + 1]]
    assert inner.tolist() == expect


def test_error_string_is_set():
    lib = L.lib()
    rc = lib.ctfp_compute_paras(None, None)
    assert rc == L.CTFP_ERR_INVALID
    assert lib.ctfp_last_error()
