"""ctypes binding of the C ABI in include/ctfp.h (libctfp.so).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C ct_forward_projection_amd/csrc``).  Loading fails loudly when it is
missing: there is no CPU fallback for the projector.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libctfp.so")

CTFP_OK = 0
CTFP_ERR_INVALID = -1
CTFP_ERR_DEVICE = -2
CTFP_ERR_NOMEM = -3
CTFP_ERR_STATE = -4
CTFP_ERR_IO = -5
CTFP_ERR_CONFIG = -6
CTFP_ERR_UNSUPPORTED = -7

CTFP_F32 = 0
CTFP_F16 = 1
CTFP_BF16 = 2

# Every symbol declared in include/ctfp.h.
EXPORTED = (
    "ctfp_last_error", "ctfp_version", "ctfp_compute_paras", "ctfp_create", "ctfp_destroy",
    "ctfp_get_system_info", "ctfp_set_volumes", "ctfp_set_labels", "ctfp_set_spectral",
    "ctfp_forward", "ctfp_line_integrals", "ctfp_flat_field", "ctfp_forward_host",
    "ctfp_postprocess", "ctfp_synchronize", "ctfp_scint_thickness", "ctfp_det_response",
    "ctfp_grid_transmission", "ctfp_grid_period_index", "ctfp_bench_copy",
    "ctfp_set_tuning", "ctfp_get_tuning", "ctfp_backward", "ctfp_set_volumes_async",
    "ctfp_get_scan", "ctfp_get_device", "ctfp_get_volumes", "ctfp_device_bytes", "ctfp_autotune",
    "ctfp_batch_create", "ctfp_batch_run", "ctfp_batch_samples", "ctfp_batch_destroy",
    "ctfp_count_samples", "ctfp_set_focal_spot", "ctfp_focal_spot_offsets",
    "ctfp_backward_noalloc", "ctfp_grid_unit_heights", "ctfp_count_marched", "ctfp_debug_fail_allocations",
)

CTFP_MARCH_GLOBAL = 0
CTFP_MARCH_LDS = 1
CTFP_ADJOINT_GATHER = 0
CTFP_ADJOINT_SCATTER = 1
CTFP_ADJOINT_CELL = 2
# bytes of the static dl/dL workspace of ctfp_backward_noalloc (per GPU)
CTFP_NOALLOC_SCRATCH_BYTES = 64 << 20


class ScanParams(C.Structure):
    """ctfp_scan_params; mirrors CTScanParas (Source/DataType.h:33-55)."""
    _fields_ = [
        ("p_num_x", C.c_int64), ("p_num_y", C.c_int64), ("p_num_z", C.c_int64),
        ("p_length_x", C.c_float), ("p_length_y", C.c_float), ("p_length_z", C.c_float),
        ("proj_num", C.c_int64),
        ("d_num_u", C.c_int64), ("d_num_v", C.c_int64),
        ("sod", C.c_float), ("sdd", C.c_float),
        ("rotated_direction", C.c_int32),
        ("d_size", C.c_float),
    ]


class SystemInfo(C.Structure):
    """ctfp_system_info; mirrors CTScanSystemInfo (Source/DataType.h:78-111)."""
    _fields_ = [
        ("rotated_direction", C.c_int32),
        ("theta_step", C.c_float),
        ("d_half_lu", C.c_float), ("d_half_lv", C.c_float),
        ("fov_r", C.c_float), ("fov_h", C.c_float),
        ("p_half_x", C.c_float), ("p_half_y", C.c_float), ("p_half_z", C.c_float),
        ("p_size_x", C.c_float), ("p_size_y", C.c_float), ("p_size_z", C.c_float),
        ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float),
        ("int_num", C.c_int64),
    ]


_FP = C.POINTER(C.c_float)


class Spectral(C.Structure):
    """ctfp_spectral."""
    _fields_ = [
        ("n_energy", C.c_int32), ("n_mat", C.c_int32),
        ("spectrum", _FP), ("atten", _FP),
        ("i0_val", C.c_float),
        ("grid", _FP), ("det", _FP),
        ("scint_mu", _FP), ("scint_thickness", _FP),
        ("det_factor", C.c_float),
        ("energy_weighting", C.c_int32),
    ]


class GridParams(C.Structure):
    """ctfp_grid_params; mirrors GridAndDetectorSystem (Source/Grid.h:42-66)."""
    _fields_ = [
        ("d_num", C.c_int64), ("d_size_um", C.c_float),
        ("interspace_width", C.c_int32), ("strip_width", C.c_int32),
        ("height_mm", C.c_float), ("focal_mm", C.c_float),
    ]


class Tuning(C.Structure):
    """ctfp_tuning: ray-march launch options (results do not depend on them)."""
    _fields_ = [("march", C.c_int32), ("chunk", C.c_int32), ("lds_bytes", C.c_int32),
                ("wave_u", C.c_int32), ("block_u", C.c_int32), ("adjoint", C.c_int32),
                ("skip_empty", C.c_int32)]


class BatchItem(C.Structure):
    """ctfp_batch_item: one scan of a batched launch."""
    _fields_ = [("ctx", C.c_void_p), ("view_begin", C.c_int64), ("view_count", C.c_int64),
                ("I", C.c_void_p), ("IAbsorb", C.c_void_p)]


class CTFPError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ctfp error {code}: {msg}")
        self.code = code


_lib = None


def _declare(lib):
    vp, i64, i32 = C.c_void_p, C.c_int64, C.c_int
    P = C.POINTER
    sig = {
        "ctfp_last_error": (C.c_char_p, []),
        "ctfp_version": (C.c_char_p, []),
        "ctfp_compute_paras": (i32, [P(ScanParams), P(SystemInfo)]),
        "ctfp_create": (i32, [P(ScanParams), i32, P(vp)]),
        "ctfp_destroy": (i32, [vp]),
        "ctfp_get_system_info": (i32, [vp, P(SystemInfo)]),
        "ctfp_set_volumes": (i32, [vp, i32, P(vp), i32, i32]),
        "ctfp_set_labels": (i32, [vp, vp, i32, i32, _FP, i32, i32]),
        "ctfp_set_spectral": (i32, [vp, P(Spectral)]),
        "ctfp_forward": (i32, [vp, i64, i64, vp, vp, vp, vp]),
        "ctfp_line_integrals": (i32, [vp, i64, i64, vp, vp]),
        "ctfp_flat_field": (i32, [vp, vp, vp]),
        "ctfp_forward_host": (i32, [vp, vp, vp, vp]),
        "ctfp_postprocess": (i32, [i64, i64, vp, vp, vp, vp, vp, vp, vp]),
        "ctfp_synchronize": (i32, [vp]),
        "ctfp_scint_thickness": (i32, [i64, C.c_float, C.c_float, _FP]),
        "ctfp_det_response": (i32, [i64, C.c_float, C.c_float, _FP, _FP]),
        "ctfp_grid_transmission": (i32, [P(GridParams), i32, _FP, _FP, _FP]),
        "ctfp_grid_period_index": (i32, [i64, _FP, P(i64), i64, P(i64)]),
        "ctfp_grid_unit_heights": (i32, [P(GridParams), _FP, i64, P(i64)]),
        "ctfp_bench_copy": (i32, [vp, vp, i64, vp]),
        "ctfp_set_tuning": (i32, [vp, P(Tuning)]),
        "ctfp_backward": (i32, [vp, i64, i64, vp, vp, vp, vp, vp]),
        "ctfp_backward_noalloc": (i32, [vp, i64, i64, vp, vp, vp, vp, vp]),
        "ctfp_get_tuning": (i32, [vp, P(Tuning)]),
        "ctfp_set_volumes_async": (i32, [vp, i32, P(vp), i32, vp]),
        "ctfp_get_scan": (i32, [vp, P(ScanParams)]),
        "ctfp_get_device": (i32, [vp, P(i32)]),
        "ctfp_get_volumes": (i32, [vp, P(i32), P(i32)]),
        "ctfp_device_bytes": (i32, [vp, P(i64)]),
        "ctfp_autotune": (i32, [vp, i32, P(Tuning)]),
        "ctfp_count_samples": (i32, [vp, i64, i64, P(i64)]),
        "ctfp_count_marched": (i32, [vp, i64, i64, P(i64)]),
        "ctfp_debug_fail_allocations": (i32, [C.c_longlong]),
        "ctfp_set_focal_spot": (i32, [vp, _FP, _FP]),
        "ctfp_focal_spot_offsets": (i32, [i64, C.c_float, C.c_uint32, _FP, _FP]),
        "ctfp_batch_create": (i32, [P(BatchItem), i32, P(vp)]),
        "ctfp_batch_run": (i32, [vp, vp]),
        "ctfp_batch_samples": (i32, [vp, P(i64)]),
        "ctfp_batch_destroy": (i32, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def lib():
    """Load libctfp.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "or `make -C ct_forward_projection_amd/csrc`")
        lib_ = C.CDLL(LIB_PATH)
        _declare(lib_)
        _lib = lib_
    return _lib


def check(rc: int) -> None:
    if rc != CTFP_OK:
        msg = lib().ctfp_last_error()
        raise CTFPError(rc, msg.decode() if msg else "")
