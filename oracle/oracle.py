"""ctypes face of libctfp_oracle.so, the CPU restatement of the reference's
method (oracle/ctfp_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libctfp_oracle.so")


class Geom(C.Structure):
    _fields_ = [("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
                ("lx", C.c_float), ("ly", C.c_float), ("lz", C.c_float),
                ("n_views", C.c_int64), ("nu", C.c_int64), ("nv", C.c_int64),
                ("sod", C.c_float), ("sdd", C.c_float), ("rot", C.c_int32),
                ("dsize", C.c_float)]


class Paras(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("theta_step", "half_u", "half_v", "fov_r", "fov_h",
                                          "half_x", "half_y", "half_z", "px", "py", "pz",
                                          "dx", "dy", "dz")] + [("n_steps", C.c_int64)]


_lib = None
_F = C.POINTER(C.c_float)
_D = C.POINTER(C.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
        l = C.CDLL(LIB_PATH)
        i64, i32, vp = C.c_int64, C.c_int, C.c_void_p
        G, P = C.POINTER(Geom), C.POINTER(Paras)
        sig = {
            "orc_compute_paras": (None, [G, P]),
            "orc_view_angle": (C.c_float, [G, P, i64]),
            "orc_march_f32": (None, [G, P, i32, _F, i64, i64, _F, i32]),
            "orc_march_f64": (None, [G, P, i32, _F, i64, i64, _D, i32]),
            "orc_sample_position_f32": (None, [G, P, i64, i64, i64, i64, _F]),
            "orc_ray_direction_f64": (None, [G, P, i64, i64, _D]),
            "orc_tex3d_f32": (C.c_float, [_F, i64, i64, i64, C.c_float, C.c_float, C.c_float]),
            "orc_tri_f64": (C.c_double, [_F, i64, i64, i64, C.c_double, C.c_double, C.c_double]),
            "orc_spectrum_f32": (None, [i64, i64, i64, i32, i32, _F, _F, _F, C.c_float, _F, _F,
                                        _F, _F, _F, i32]),
            "orc_spectrum_f64": (None, [i64, i64, i64, i32, i32, _D, _F, _F, C.c_double, _F, _F,
                                        _D, _D, _D, i32]),
            "orc_forward_literal_f32": (None, [G, P, vp, i32, _F, i32, _F, _F, C.c_float, _F, _F,
                                               _F, _F, _F, i32]),
            "orc_labels_to_density_channels": (None, [vp, i64, i32, _F, _F]),
            "orc_labels_to_density_single": (None, [vp, i64, C.c_float, _F]),
            "orc_msvc_rand_sequence": (None, [C.c_uint32, i32, C.POINTER(C.c_int)]),
            "orc_scint_thickness": (None, [i64, C.c_float, C.c_float, _F]),
            "orc_det_response": (None, [i64, C.c_float, C.c_float, _F, _F]),
            "orc_grid_transmission": (i32, [i64, i64, i32, i32, C.c_float, C.c_float, C.c_float,
                                            C.c_float, _F]),
            "orc_grid_unit_heights": (None, [i64, C.c_float, C.c_float, _F]),
            "orc_postprocess": (None, [i64, i64, _F, _F, _F, vp, vp, _F]),
            "orc_num_threads": (i32, []),
            "orc_backproject_f64": (None, [G, P, i32, i64, i64, _D, _D]),
            "orc_march_focal_f64": (None, [G, P, i32, _F, i64, i64, _D, _D, i32]),
            "orc_backproject_focal_f64": (None, [G, P, i32, i64, i64, _D, _D, _D]),
            "orc_backproject_box_f64": (None, [G, P, i32, i64, i64, _D, _D, C.POINTER(i64),
                                               C.POINTER(i64), _D]),
            "orc_focal_offsets": (None, [i64, C.c_float, C.c_uint32, _F, _F]),
            "orc_spectrum_grad_f64": (None, [i64, i64, i64, i32, i32, _D, _F, _F, C.c_double, _F, _F,
                                             _D, _D, _D]),
        }
        for name, (res, args) in sig.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def _f(a):
    return None if a is None else a.ctypes.data_as(_F)


def _d(a):
    return None if a is None else a.ctypes.data_as(_D)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def geom(nx, ny, nz, lx, ly, lz, n_views, nu, nv, sod, sdd, dsize, rot=1) -> Geom:
    return Geom(nx, ny, nz, lx, ly, lz, n_views, nu, nv, sod, sdd, rot, dsize)


def geom_from_setup(setup, views=None) -> Geom:
    return geom(*setup.n_vox, *setup.length_mm, setup.views if views is None else views,
                *setup.det, setup.sod, setup.sdd, setup.d_size, setup.rot)


def paras(g: Geom) -> Paras:
    p = Paras()
    lib().orc_compute_paras(C.byref(g), C.byref(p))
    return p


def view_angle(g, p, view):
    return lib().orc_view_angle(C.byref(g), C.byref(p), view)


def march_f32(g, p, vols, view0=0, n_views=None, threads=0):
    vols = f32(np.stack([np.asarray(v) for v in vols]))
    n_mat = vols.shape[0]
    n_views = g.n_views - view0 if n_views is None else n_views
    out = np.empty((n_views, n_mat, g.nv, g.nu), dtype=np.float32)
    lib().orc_march_f32(C.byref(g), C.byref(p), n_mat, _f(vols), view0, n_views, _f(out), threads)
    return out


def march_f64(g, p, vols, view0=0, n_views=None, threads=0, focal=None):
    """FP64 line integrals [views][M][nv][nu]; focal: per-view source offsets
    [all views][2] in mm (detector u and v directions) or None."""
    vols = f32(np.stack([np.asarray(v) for v in vols]))
    n_mat = vols.shape[0]
    n_views = g.n_views - view0 if n_views is None else n_views
    out = np.empty((n_views, n_mat, g.nv, g.nu), dtype=np.float64)
    if focal is None:
        lib().orc_march_f64(C.byref(g), C.byref(p), n_mat, _f(vols), view0, n_views, _d(out), threads)
    else:
        fo = np.ascontiguousarray(focal, dtype=np.float64)
        lib().orc_march_focal_f64(C.byref(g), C.byref(p), n_mat, _f(vols), view0, n_views, _d(fo),
                                  _d(out), threads)
    return out


def focal_offsets(n_views, spot_mm, seed=10):
    """Per-view focal-spot offsets (u, v) in mm: the reference's generator
    (CBCTForwardProj.cpp:122-148) with the disc test u^2 + v^2 <= r^2."""
    ou = np.empty(n_views, np.float32)
    ov = np.empty(n_views, np.float32)
    lib().orc_focal_offsets(n_views, spot_mm, seed, _f(ou), _f(ov))
    return np.stack([ou, ov], axis=1)


def spectrum_f32(L, atten, spectrum, i0val, grid=None, det=None, threads=0):
    L = f32(L)
    nv_, n_mat, nv, nu = L.shape
    atten, spectrum = f32(atten), f32(spectrum)
    grid = None if grid is None else f32(grid)
    det = None if det is None else f32(det)
    I = np.empty((nv_, nv, nu), dtype=np.float32)
    IA = np.empty_like(I)
    I0 = np.empty((nv, nu), dtype=np.float32)
    lib().orc_spectrum_f32(nu, nv, nv_, n_mat, spectrum.size, _f(L), _f(atten), _f(spectrum),
                           i0val, _f(grid), _f(det), _f(I), _f(IA), _f(I0), threads)
    return I, IA, I0


def spectrum_f64(L, atten, spectrum, i0val, grid=None, det=None, threads=0):
    L = np.ascontiguousarray(L, dtype=np.float64)
    nv_, n_mat, nv, nu = L.shape
    atten, spectrum = f32(atten), f32(spectrum)
    grid = None if grid is None else f32(grid)
    det = None if det is None else f32(det)
    I = np.empty((nv_, nv, nu), dtype=np.float64)
    IA = np.empty_like(I)
    I0 = np.empty((nv, nu), dtype=np.float64)
    lib().orc_spectrum_f64(nu, nv, nv_, n_mat, spectrum.size, _d(L), _f(atten), _f(spectrum),
                           float(i0val), _f(grid), _f(det), _d(I), _d(IA), _d(I0), threads)
    return I, IA, I0


def forward_literal_f32(g, p, labels, density, mass_atten, spectrum, i0val, grid=None, det=None,
                        threads=0):
    labels = np.ascontiguousarray(labels, dtype=np.uint8)
    density, mass_atten, spectrum = f32(density), f32(mass_atten), f32(spectrum)
    grid = None if grid is None else f32(grid)
    det = None if det is None else f32(det)
    I = np.empty((g.n_views, g.nv, g.nu), dtype=np.float32)
    IA = np.empty_like(I)
    I0 = np.empty((g.nv, g.nu), dtype=np.float32)
    lib().orc_forward_literal_f32(C.byref(g), C.byref(p), labels.ctypes.data, density.size,
                                  _f(density), spectrum.size, _f(mass_atten), _f(spectrum), i0val,
                                  _f(grid), _f(det), _f(I), _f(IA), _f(I0), threads)
    return I, IA, I0


def labels_to_density_channels(labels, density):
    labels = np.ascontiguousarray(labels, dtype=np.uint8)
    density = f32(density)
    out = np.empty((density.size,) + labels.shape, dtype=np.float32)
    lib().orc_labels_to_density_channels(labels.ctypes.data, labels.size, density.size,
                                         _f(density), _f(out))
    return out


def labels_to_density_single(labels, rho0):
    labels = np.ascontiguousarray(labels, dtype=np.uint8)
    out = np.empty(labels.shape, dtype=np.float32)
    lib().orc_labels_to_density_single(labels.ctypes.data, labels.size, rho0, _f(out))
    return out


def msvc_rand_sequence(seed, n):
    out = np.empty(n, dtype=np.int32)
    lib().orc_msvc_rand_sequence(seed, n, out.ctypes.data_as(C.POINTER(C.c_int)))
    return out


def scint_thickness(npix, t, terr):
    out = np.empty(npix, dtype=np.float32)
    lib().orc_scint_thickness(npix, t, terr, _f(out))
    return out


def det_response(factor, mu_c, thickness):
    t = f32(thickness)
    out = np.empty_like(t)
    lib().orc_det_response(t.size, factor, mu_c, _f(t), _f(out))
    return out


def grid_transmission(d_num, d_size_um, w, sw, h_mm, fd_mm, mu_s, mu_i):
    out = np.empty(d_num, dtype=np.float32)
    rc = lib().orc_grid_transmission(d_num, d_size_um, w, sw, h_mm, fd_mm, mu_s, mu_i, _f(out))
    if rc != 0:
        raise ValueError("degenerate grid layout")
    return out


def grid_unit_heights(n_units, h_mm, fd_mm):
    out = np.empty(n_units, dtype=np.float32)
    lib().orc_grid_unit_heights(n_units, h_mm, fd_mm, _f(out))
    return out


def tex3d_f32(vol, x, y, z):
    vol = f32(vol)
    nz, ny, nx = vol.shape
    return lib().orc_tex3d_f32(_f(vol), nx, ny, nz, x, y, z)


def tri_f64(vol, x, y, z):
    vol = f32(vol)
    nz, ny, nx = vol.shape
    return lib().orc_tri_f64(_f(vol), nx, ny, nz, x, y, z)


def sample_position_f32(g, p, view, u, v, k):
    out = np.empty(3, dtype=np.float32)
    lib().orc_sample_position_f32(C.byref(g), C.byref(p), view, u, v, k, _f(out))
    return out


def ray_direction_f64(g, p, u, v):
    out = np.empty(3, dtype=np.float64)
    lib().orc_ray_direction_f64(C.byref(g), C.byref(p), u, v, _d(out))
    return out


def postprocess(I, IA, I0):
    I, IA, I0 = f32(I), f32(IA), f32(I0)
    n_views = I.shape[0]
    npix = I0.size
    Iu = np.empty(I.shape, dtype=np.uint32)
    IAu = np.empty(I.shape, dtype=np.uint32)
    proj = np.empty(I.shape, dtype=np.float32)
    lib().orc_postprocess(n_views, npix, _f(I), _f(IA), _f(I0), Iu.ctypes.data, IAu.ctypes.data,
                          _f(proj))
    return Iu, IAu, proj


def num_threads():
    return lib().orc_num_threads()


def backproject_f64(g, p, gL, view0=0, focal=None):
    """Adjoint of march_f64: gL [views][M][nv][nu] -> grad [M][nz][ny][nx]."""
    gL = np.ascontiguousarray(gL, dtype=np.float64)
    n_views, n_mat = gL.shape[:2]
    grad = np.zeros((n_mat, g.nz, g.ny, g.nx), dtype=np.float64)
    if focal is None:
        lib().orc_backproject_f64(C.byref(g), C.byref(p), n_mat, view0, n_views, _d(gL), _d(grad))
    else:
        fo = np.ascontiguousarray(focal, dtype=np.float64)
        lib().orc_backproject_focal_f64(C.byref(g), C.byref(p), n_mat, view0, n_views, _d(fo), _d(gL),
                                        _d(grad))
    return grad


def backproject_box_f64(g, p, gL, box0, boxn, view0=0):
    """backproject_f64 into the voxel box [box0, box0 + boxn) ((z, y, x)
    origin and extent) only: grad [M][boxn_z][boxn_y][boxn_x]."""
    gL = np.ascontiguousarray(gL, dtype=np.float64)
    n_views, n_mat = gL.shape[:2]
    b0 = (C.c_int64 * 3)(*[int(v) for v in box0])
    bn = (C.c_int64 * 3)(*[int(v) for v in boxn])
    grad = np.zeros((n_mat,) + tuple(int(v) for v in boxn), dtype=np.float64)
    lib().orc_backproject_box_f64(C.byref(g), C.byref(p), n_mat, view0, n_views, None, _d(gL), b0, bn, _d(grad))
    return grad


def spectrum_grad_f64(L, atten, spectrum, i0val, gI, gIA, grid=None, det=None):
    """dl/dL [views][M][nv][nu] from dl/dI and dl/dIA (either may be None)."""
    L = np.ascontiguousarray(L, dtype=np.float64)
    nv_, n_mat, nv, nu = L.shape
    atten, spectrum = f32(atten), f32(spectrum)
    grid = None if grid is None else f32(grid)
    det = None if det is None else f32(det)
    gI = None if gI is None else np.ascontiguousarray(gI, dtype=np.float64)
    gIA = None if gIA is None else np.ascontiguousarray(gIA, dtype=np.float64)
    gL = np.empty_like(L)
    lib().orc_spectrum_grad_f64(nu, nv, nv_, n_mat, spectrum.size, _d(L), _f(atten), _f(spectrum),
                                float(i0val), _f(grid), _f(det), _d(gI), _d(gIA), _d(gL))
    return gL
