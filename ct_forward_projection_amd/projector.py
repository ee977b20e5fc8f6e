"""Python face of the C ABI: scan set-up, the resident projector context, and
the host-side detector / grid models.

Device buffers are torch tensors (torch is the allocator and stream provider
here); every computation runs in libctfp.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def scan_params(**kw) -> L.ScanParams:
    """ctfp_scan_params from keyword arguments (the member names)."""
    p = L.ScanParams()
    defaults = dict(rotated_direction=1)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(p, k, v)
    return p


def compute_paras(scan: L.ScanParams) -> dict:
    """CBCTForwardProj::computeParas (Source/CBCTForwardProj.cpp:23-62)."""
    out = L.SystemInfo()
    L.check(L.lib().ctfp_compute_paras(C.byref(scan), C.byref(out)))
    return {name: getattr(out, name) for name, _ in L.SystemInfo._fields_}


# ------------------------------------------------------------ host models

def scint_thickness(npix: int, thickness: float, thickness_err: float) -> np.ndarray:
    """Scintillator thickness map (CBCTForwardProj.cpp:99-111)."""
    out = np.empty(npix, dtype=np.float32)
    L.check(L.lib().ctfp_scint_thickness(npix, thickness, thickness_err, _fptr(out)))
    return out


def det_response(factor: float, mu_c: float, thickness: np.ndarray) -> np.ndarray:
    """eta = f (1 - exp(-mu_c t)) (CBCTForwardProj.cpp:114-120)."""
    t = np.ascontiguousarray(thickness, dtype=np.float32)
    out = np.empty_like(t)
    L.check(L.lib().ctfp_det_response(t.size, factor, mu_c, _fptr(t), _fptr(out)))
    return out


def grid_transmission(d_num: int, d_size_mm: float, interspace_width: int, strip_width: int,
                      height_mm: float, focal_mm: float, mu_strip, mu_inter) -> np.ndarray:
    """T_E(u) [n_energy][d_num] of the focused grid (Grid.cpp:373-443).  The
    pitch enters as FP32 dSize * 1000 um (CBCTPolyForwardProjGrid.cpp:291)."""
    d_um = float(np.float32(d_size_mm) * np.float32(1000))
    g = L.GridParams(d_num, d_um, interspace_width, strip_width, height_mm, focal_mm)
    ms = np.ascontiguousarray(np.atleast_1d(mu_strip), dtype=np.float32)
    mi = np.ascontiguousarray(np.atleast_1d(mu_inter), dtype=np.float32)
    if ms.shape != mi.shape:
        raise ValueError("mu_strip and mu_inter must have the same length")
    T = np.empty((ms.size, d_num), dtype=np.float32)
    L.check(L.lib().ctfp_grid_transmission(C.byref(g), ms.size, _fptr(ms), _fptr(mi), _fptr(T)))
    return T


def grid_unit_heights(d_num: int, d_size_mm: float, interspace_width: int, strip_width: int,
                      height_mm: float, focal_mm: float) -> np.ndarray:
    """Slant heights (um) of the grid's 1-um lamella units (Grid.cpp:76-87)."""
    d_um = float(np.float32(d_size_mm) * np.float32(1000))
    g = L.GridParams(d_num, d_um, interspace_width, strip_width, height_mm, focal_mm)
    n = C.c_int64()
    L.check(L.lib().ctfp_grid_unit_heights(C.byref(g), None, 0, C.byref(n)))
    H = np.empty(n.value, dtype=np.float32)
    L.check(L.lib().ctfp_grid_unit_heights(C.byref(g), _fptr(H), H.size, C.byref(n)))
    return H


def focal_spot_offsets(n_views: int, spot_mm: float, seed: int = 10) -> np.ndarray:
    """Per-view focal-spot offsets [n_views][2] in mm (CBCTForwardProj.cpp:122-148,
    with the disc test u^2 + v^2 <= r^2)."""
    ou = np.empty(n_views, dtype=np.float32)
    ov = np.empty(n_views, dtype=np.float32)
    L.check(L.lib().ctfp_focal_spot_offsets(n_views, spot_mm, seed, _fptr(ou), _fptr(ov)))
    return np.stack([ou, ov], axis=1)


def grid_period_index(T: np.ndarray) -> np.ndarray:
    """Grid.cpp:605-620: 0, strict local minima, d_num-1."""
    t = np.ascontiguousarray(T, dtype=np.float32)
    cap = t.size + 2
    out = np.empty(cap, dtype=np.int64)
    n = C.c_int64(0)
    L.check(L.lib().ctfp_grid_period_index(t.size, _fptr(t),
                                           out.ctypes.data_as(C.POINTER(C.c_int64)), cap,
                                           C.byref(n)))
    return out[: n.value].copy()


# ------------------------------------------------------------- projector

def _torch():
    import torch
    return torch


def _stream_handle(stream):
    if stream is None:
        return _torch().cuda.current_stream().cuda_stream
    return getattr(stream, "cuda_stream", stream)


class Projector:
    """One scan resident on one GPU: volumes, spectral weights, geometry.

    Mirrors the life cycle of the reference's device layer (initDevice* ->
    per-energy forward*GPU -> freeDeviceMemory, CBCTPolyProjGPU.cuh:33-64) with
    the energy loop folded into one call.
    """

    def __init__(self, scan: L.ScanParams, device: int = 0):
        self.scan = scan
        self.device = device
        self.nu, self.nv, self.views = scan.d_num_u, scan.d_num_v, scan.proj_num
        self.n_mat = 0
        self._ctx = C.c_void_p()
        self._graphs = 0            # autograd graphs that still use the context
        self._close_pending = False  # close() waits for them
        L.check(L.lib().ctfp_create(C.byref(scan), device, C.byref(self._ctx)))
        info = L.SystemInfo()
        L.check(L.lib().ctfp_get_system_info(self._live(), C.byref(info)))
        self.info = {name: getattr(info, name) for name, _ in L.SystemInfo._fields_}

    # -- lifetime
    def close(self):
        """Releases the context.  While an autograd graph built by
        autograd.project still uses it, the projector is only marked closed
        (no new calls) and the context is released with the last graph."""
        if not self._ctx:
            return
        if self._graphs > 0:
            self._close_pending = True
            return
        self._destroy()

    def _live(self):
        """The context pointer; raises once the projector is closed."""
        if not self._ctx or self._close_pending:
            raise L.CTFPError(L.CTFP_ERR_STATE, "the projector is closed")
        return self._ctx

    def _destroy(self):
        L.lib().ctfp_destroy(self._ctx)
        self._ctx = C.c_void_p()

    def _graph_token(self):
        """A token that an autograd graph holds while it uses this context;
        the context is not released before every token is gone."""
        import weakref

        class _Token:
            pass

        tok = _Token()
        self._graphs += 1
        weakref.finalize(tok, self._graph_released)
        return tok

    def _graph_released(self):
        self._graphs -= 1
        if self._graphs == 0 and self._close_pending and self._ctx:
            self._destroy()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_tuning(self, march: str | None = None, chunk: int | None = None,
                   lds_bytes: int | None = None, wave_u: int | None = None,
                   block_u: int | None = None, adjoint: str | None = None,
                   skip_empty: bool | None = None):
        """Launch options of the march ("global" or "lds", samples per staged
        chunk, LDS bytes per workgroup, ray-tile shape, skipping of chunks
        whose cells are all zero; projections are identical for all) and of
        the adjoint ("gather", "cell" or "scatter")."""
        t = L.Tuning()
        L.check(L.lib().ctfp_get_tuning(self._live(), C.byref(t)))
        if march is not None:
            t.march = {"global": L.CTFP_MARCH_GLOBAL, "lds": L.CTFP_MARCH_LDS}[march]
        if adjoint is not None:
            t.adjoint = {"gather": L.CTFP_ADJOINT_GATHER, "scatter": L.CTFP_ADJOINT_SCATTER,
                         "cell": L.CTFP_ADJOINT_CELL}[adjoint]
        if skip_empty is not None:
            t.skip_empty = int(bool(skip_empty))
        for name, v in (("chunk", chunk), ("lds_bytes", lds_bytes), ("wave_u", wave_u),
                        ("block_u", block_u)):
            if v is not None:
                setattr(t, name, v)
        L.check(L.lib().ctfp_set_tuning(self._live(), C.byref(t)))

    def autotune(self, max_lds_bytes: int = 0) -> dict:
        """Time launch options on a sample of this scan's views, keep the
        fastest (results do not depend on it) and return it.  max_lds_bytes
        caps the LDS budget (for scans that share a batched launch)."""
        t = L.Tuning()
        L.check(L.lib().ctfp_autotune(self._live(), int(max_lds_bytes), C.byref(t)))
        return self.get_tuning()

    @property
    def handle(self) -> int:
        """The ctfp_context pointer as an integer (for the PyTorch extension)."""
        if not self._ctx or self._close_pending:
            raise L.CTFPError(L.CTFP_ERR_STATE, "the projector is closed")
        return int(self._ctx.value)

    @property
    def device_bytes(self) -> int:
        """Device memory held by the projector itself (volume layout, tables)."""
        b = C.c_int64()
        L.check(L.lib().ctfp_device_bytes(self._live(), C.byref(b)))
        return int(b.value)

    def get_tuning(self) -> dict:
        t = L.Tuning()
        L.check(L.lib().ctfp_get_tuning(self._live(), C.byref(t)))
        return {"march": "lds" if t.march == L.CTFP_MARCH_LDS else "global", "chunk": t.chunk,
                "lds_bytes": t.lds_bytes, "wave_u": t.wave_u, "block_u": t.block_u,
                "adjoint": {L.CTFP_ADJOINT_SCATTER: "scatter", L.CTFP_ADJOINT_CELL: "cell"}.get(t.adjoint, "gather"),
                "skip_empty": bool(t.skip_empty)}

    def set_focal_spot(self, offsets=None):
        """Per-view focal-spot source offsets [proj_num][2] (mm, along the
        detector's u and v), or None for a point source."""
        if offsets is None:
            L.check(L.lib().ctfp_set_focal_spot(self._live(), None, None))
            return
        o = np.ascontiguousarray(offsets, dtype=np.float32)
        if o.shape != (self.views, 2):
            raise ValueError(f"focal offsets must be [{self.views}][2]")
        ou, ov = np.ascontiguousarray(o[:, 0]), np.ascontiguousarray(o[:, 1])
        L.check(L.lib().ctfp_set_focal_spot(self._live(), _fptr(ou), _fptr(ov)))

    def marched_samples(self, view_begin: int = 0, view_count: int | None = None) -> int:
        """Samples the march evaluates (rays' windows clipped to the volume)."""
        vb, vc = self._views(view_begin, view_count)
        n = C.c_int64()
        L.check(L.lib().ctfp_count_samples(self._live(), vb, vc, C.byref(n)))
        return int(n.value)

    def evaluated_samples(self, view_begin: int = 0, view_count: int | None = None) -> int:
        """Samples the LDS march evaluates: the clipped windows less the
        chunks skipped as empty (with the current tuning)."""
        vb, vc = self._views(view_begin, view_count)
        n = C.c_int64()
        L.check(L.lib().ctfp_count_marched(self._live(), vb, vc, C.byref(n)))
        return int(n.value)

    @property
    def samples_per_view(self) -> int:
        return int(self.nu * self.nv * self.info["int_num"])

    # -- inputs
    def set_volumes(self, volumes, storage_dtype: int = L.CTFP_F32):
        """Density volumes [z][y][x] FP32: numpy arrays (host) or CUDA tensors."""
        vols = list(volumes)
        on_device = not isinstance(vols[0], np.ndarray)
        keep = []
        ptrs = (C.c_void_p * len(vols))()
        for i, v in enumerate(vols):
            if on_device:
                v = v.contiguous()
                if v.dtype != _torch().float32:
                    raise TypeError("volumes must be float32")
                ptrs[i] = v.data_ptr()
            else:
                v = np.ascontiguousarray(v, dtype=np.float32)
                ptrs[i] = v.ctypes.data
            expect = (self.scan.p_num_z, self.scan.p_num_y, self.scan.p_num_x)
            if tuple(v.shape) != expect:
                raise ValueError(f"volume {i} has shape {tuple(v.shape)}, expected {expect}")
            keep.append(v)
        L.check(L.lib().ctfp_set_volumes(self._live(), len(vols), ptrs, int(on_device), storage_dtype))
        self.n_mat = len(vols)

    def set_labels(self, labels: np.ndarray, density, single_material: bool,
                   storage_dtype: int = L.CTFP_F32):
        lab = np.ascontiguousarray(labels, dtype=np.uint8)
        rho = np.ascontiguousarray(np.atleast_1d(density), dtype=np.float32)
        n_mat = rho.size
        L.check(L.lib().ctfp_set_labels(self._live(), lab.ctypes.data, 0, n_mat, _fptr(rho),
                                        int(single_material), storage_dtype))
        self.n_mat = 1 if single_material else n_mat

    def set_spectral(self, spectrum, atten, i0_val: float, grid=None, det=None,
                     scint_mu=None, scint_thickness=None, det_factor: float = 1.0,
                     energy_weighting: bool = False):
        """spectrum [E]; atten [E][n_mat] (1/mm per g/cm^3); grid [E][nu];
        det [E][nv*nu]; or scint_mu [E] with scint_thickness [nv*nu]."""
        s = np.ascontiguousarray(spectrum, dtype=np.float32)
        a = np.ascontiguousarray(atten, dtype=np.float32).reshape(s.size, -1)
        keep = [s, a]
        sp = L.Spectral()
        sp.n_energy = s.size
        sp.n_mat = a.shape[1]
        sp.spectrum = _fptr(s)
        sp.atten = _fptr(a)
        sp.i0_val = i0_val
        sp.det_factor = det_factor
        sp.energy_weighting = int(energy_weighting)
        if grid is not None:
            g = np.ascontiguousarray(grid, dtype=np.float32).reshape(s.size, self.nu)
            keep.append(g)
            sp.grid = _fptr(g)
        if det is not None:
            d = np.ascontiguousarray(det, dtype=np.float32).reshape(s.size, self.nu * self.nv)
            keep.append(d)
            sp.det = _fptr(d)
        elif scint_mu is not None:
            m = np.ascontiguousarray(scint_mu, dtype=np.float32).reshape(s.size)
            t = np.ascontiguousarray(scint_thickness, dtype=np.float32).reshape(self.nu * self.nv)
            keep += [m, t]
            sp.scint_mu = _fptr(m)
            sp.scint_thickness = _fptr(t)
        L.check(L.lib().ctfp_set_spectral(self._live(), C.byref(sp)))

    # -- compute (device)
    def _views(self, view_begin, view_count):
        if view_count is None:
            view_count = self.views - view_begin
        return int(view_begin), int(view_count)

    def forward(self, view_begin: int = 0, view_count: int | None = None, I=None, IA=None,
                L_out=None, want_L: bool = False, stream=None):
        """Polyenergetic projection; returns (I, IAbsorb[, L]) CUDA tensors."""
        torch = _torch()
        vb, vc = self._views(view_begin, view_count)
        dev = torch.device("cuda", self.device)
        if I is None:
            I = torch.empty((vc, self.nv, self.nu), dtype=torch.float32, device=dev)
        if IA is None:
            IA = torch.empty((vc, self.nv, self.nu), dtype=torch.float32, device=dev)
        if want_L and L_out is None:
            L_out = torch.empty((vc, self.n_mat, self.nv, self.nu), dtype=torch.float32, device=dev)
        for t in (I, IA) + ((L_out,) if L_out is not None else ()):
            if not t.is_contiguous() or t.dtype != torch.float32:
                raise ValueError("outputs must be contiguous float32 CUDA tensors")
        if I.numel() != vc * self.nv * self.nu or IA.numel() != I.numel():
            raise ValueError("output size does not match the view range")
        if L_out is not None and L_out.numel() != vc * self.n_mat * self.nv * self.nu:
            raise ValueError("L size does not match the view range")
        L.check(L.lib().ctfp_forward(self._live(), vb, vc, I.data_ptr(), IA.data_ptr(),
                                     L_out.data_ptr() if L_out is not None else None,
                                     _stream_handle(stream)))
        return (I, IA, L_out) if want_L or L_out is not None else (I, IA)

    def line_integrals(self, view_begin: int = 0, view_count: int | None = None, out=None,
                       stream=None):
        torch = _torch()
        vb, vc = self._views(view_begin, view_count)
        if out is None:
            out = torch.empty((vc, self.n_mat, self.nv, self.nu), dtype=torch.float32,
                              device=torch.device("cuda", self.device))
        if out.numel() != vc * self.n_mat * self.nv * self.nu or not out.is_contiguous():
            raise ValueError("L output has the wrong size or layout")
        L.check(L.lib().ctfp_line_integrals(self._live(), vb, vc, out.data_ptr(),
                                            _stream_handle(stream)))
        return out

    def backward(self, gI=None, gIA=None, L_in=None, view_begin: int = 0,
                 view_count: int | None = None, grad=None, stream=None):
        """Adjoint of forward: adds dl/d(density) for upstream gradients
        gI = dl/dI and gIA = dl/dIAbsorb ([views][nv][nu] CUDA tensors, either
        may be None) to `grad` ([n_mat][nz][ny][nx], zero-initialised when not
        given) and returns it.  L_in (from forward(want_L=True)) skips the
        per-ray recomputation of the line integrals."""
        torch = _torch()
        vb, vc = self._views(view_begin, view_count)
        if grad is None:
            grad = torch.zeros((self.n_mat, self.scan.p_num_z, self.scan.p_num_y, self.scan.p_num_x),
                               dtype=torch.float32, device=torch.device("cuda", self.device))
        shape = (vc, self.nv, self.nu)
        for t in (gI, gIA):
            if t is not None and (tuple(t.shape) != shape or t.dtype != torch.float32
                                  or not t.is_contiguous()):
                raise ValueError(f"upstream gradients must be contiguous float32 {shape}")
        if L_in is not None and L_in.numel() != vc * self.n_mat * self.nv * self.nu:
            raise ValueError("L_in does not match the view range")
        if grad.numel() != self.n_mat * self.scan.p_num_x * self.scan.p_num_y * self.scan.p_num_z \
                or not grad.is_contiguous() or grad.dtype != torch.float32:
            raise ValueError("grad must be a contiguous float32 [n_mat][nz][ny][nx] tensor")
        ptr = lambda t: None if t is None else t.data_ptr()
        L.check(L.lib().ctfp_backward(self._live(), vb, vc, ptr(gI), ptr(gIA), ptr(L_in), grad.data_ptr(),
                                      _stream_handle(stream)))
        return grad

    def flat_field(self, stream=None):
        torch = _torch()
        out = torch.empty((self.nv, self.nu), dtype=torch.float32,
                          device=torch.device("cuda", self.device))
        L.check(L.lib().ctfp_flat_field(self._live(), out.data_ptr(), _stream_handle(stream)))
        return out

    # -- compute (host buffers)
    def forward_host(self):
        """Whole scan into numpy arrays: (I, IAbsorb, I0)."""
        I = np.empty((self.views, self.nv, self.nu), dtype=np.float32)
        IA = np.empty_like(I)
        I0 = np.empty((self.nv, self.nu), dtype=np.float32)
        L.check(L.lib().ctfp_forward_host(self._live(), I.ctypes.data, IA.ctypes.data, I0.ctypes.data))
        return I, IA, I0


class Batch:
    """Several scans in one persistent launch, captured once into a HIP graph.

    `items` is a list of (projector, view_begin, view_count, I, IAbsorb) with
    I / IAbsorb contiguous float32 CUDA tensors [view_count][nv][nu].  The
    projectors (one device, one storage type, LDS march) must keep their
    volumes and spectral tables while the batch exists."""

    def __init__(self, items):
        torch = _torch()
        arr = (L.BatchItem * len(items))()
        self._keep = []
        for i, (proj, vb, vc, I, IA) in enumerate(items):
            shape = (vc, proj.nv, proj.nu)
            for t in (I, IA):
                if tuple(t.shape) != shape or t.dtype != torch.float32 or not t.is_contiguous():
                    raise ValueError(f"batch item {i}: outputs must be contiguous float32 {shape}")
            arr[i] = L.BatchItem(proj._live().value, vb, vc, I.data_ptr(), IA.data_ptr())
            self._keep.append((proj, I, IA))
        self._b = C.c_void_p()
        L.check(L.lib().ctfp_batch_create(arr, len(items), C.byref(self._b)))

    def run(self, stream=None):
        L.check(L.lib().ctfp_batch_run(self._b, _stream_handle(stream)))

    @property
    def samples(self) -> int:
        n = C.c_int64()
        L.check(L.lib().ctfp_batch_samples(self._b, C.byref(n)))
        return int(n.value)

    def close(self):
        if self._b:
            L.lib().ctfp_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def postprocess(I, IA, I0, stream=None):
    """Floor to uint32 and p = ln(I0 / floor(I)) on the device
    (CBCTForwardProj.cpp:79-90, 376-387).  Returns (I_u32, IA_u32, proj);
    the uint32 images come back as int32 tensors holding the same bits."""
    torch = _torch()
    n_views = I.shape[0]
    npix = I[0].numel()
    Iu = torch.empty(I.shape, dtype=torch.int32, device=I.device)
    IAu = torch.empty(I.shape, dtype=torch.int32, device=I.device)
    proj = torch.empty(I.shape, dtype=torch.float32, device=I.device)
    L.check(L.lib().ctfp_postprocess(n_views, npix, I.data_ptr(), IA.data_ptr(), I0.data_ptr(),
                                     Iu.data_ptr(), IAu.data_ptr(), proj.data_ptr(),
                                     _stream_handle(stream)))
    return Iu, IAu, proj


def spectral_inputs(setup, nu: int, nv: int, with_response: bool = True,
                    with_grid: bool | None = None) -> dict:
    """Host tables of a synthetic ScanSetup, ready for Projector.set_spectral."""
    spec = setup.spectrum()
    kw = dict(spectrum=spec, atten=setup.atten(), i0_val=setup.i0_val,
              energy_weighting=setup.energy_weighting, det_factor=setup.det_factor)
    use_grid = (setup.grid is not None) if with_grid is None else with_grid
    if use_grid:
        g = setup.grid
        mu_s, mu_i = setup.grid_mu()
        kw["grid"] = grid_transmission(nu, setup.d_size, g["interspace_width"], g["strip_width"],
                                       g["height_mm"], g["focal_mm"], mu_s, mu_i)
    if with_response:
        kw["scint_mu"] = setup.scint_mu()
        kw["scint_thickness"] = scint_thickness(nu * nv, setup.scint_thickness,
                                                setup.scint_thickness_err)
    return kw
