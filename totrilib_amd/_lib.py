"""ctypes binding of libtotri_hip.so (the C ABI in include/totri_hip.h).

The library is built in-tree by totrilib_amd.build.  Importing this module
never falls back to another implementation: if the shared library is missing
or does not load, every call raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtotri_hip.so")

TOTRI_OK = 0
TOTRI_EINVAL = 1
TOTRI_EHIP = 2
TOTRI_EWORKSPACE = 3

TOTRI_F32 = 0
TOTRI_BF16 = 1
TOTRI_F16 = 2

TOTRI_FK_SHIFTED = 0
TOTRI_FK_UNSHIFTED = 1
TOTRI_FK_UNSHIFTED_ONESIDED = 2

_c_i64 = ctypes.c_int64
_c_p = ctypes.c_void_p
_c_f = ctypes.c_float
_c_sz = ctypes.c_size_t
_c_int = ctypes.c_int
_F3 = ctypes.c_float * 3

# name -> (restype, argtypes)
SIGNATURES: dict[str, tuple] = {
    "totri_last_error": (ctypes.c_char_p, []),
    "totri_abi_version": (_c_int, []),
    "totri_volume_confocal_forward_workspace": (_c_sz, [_c_i64] * 5),
    "totri_volume_confocal_forward": (
        _c_int,
        [_c_p, _c_i64, _c_i64, _c_i64, _c_p, _c_i64, _c_p, _c_i64, _c_int,
         _F3, _F3, _c_f, _c_f, _c_p, _c_sz, _c_p]),
    "totri_volume_confocal_grad_volume_workspace": (_c_sz, [_c_i64] * 5),
    "totri_volume_confocal_grad_volume": (
        _c_int,
        [_c_p, _c_i64, _c_i64, _c_i64, _c_p, _c_i64, _c_p, _c_i64, _c_int,
         _F3, _F3, _c_f, _c_f, _c_p, _c_sz, _c_p]),
    "totri_volume_confocal_grad_volume_window": (
        _c_int,
        [_c_p, _c_i64, _c_i64, _c_i64, _c_p, _c_i64, _c_p, _c_i64, _c_int,
         _F3, _F3, _c_f, _c_f, _c_p]),
    "totri_volume_confocal_forward_f64": (
        _c_int,
        [_c_p, _c_i64, _c_i64, _c_i64, _c_p, _c_i64, _c_p, _c_i64, _F3, _F3, _c_f, _c_f, _c_p]),
    "totri_volume_confocal_grad_volume_f64": (
        _c_int,
        [_c_p, _c_i64, _c_i64, _c_i64, _c_p, _c_i64, _c_p, _c_i64, _F3, _F3, _c_f, _c_f, _c_int,
         _c_p]),
    "totri_backprojection_confocal_velten_forward_f64": (
        _c_int,
        [_c_p, _c_i64, _c_i64, _c_i64, _c_p, _c_p, _c_i64, _c_p, _c_i64, _c_i64, _c_i64,
         _F3, _F3, _c_f, _c_f, _c_p]),
    "totri_volume_confocal_grad_measurement_points_workspace": (_c_sz, [_c_i64] * 5),
    "totri_volume_confocal_grad_measurement_points": (
        _c_int,
        [_c_p, _c_p, _c_i64, _c_i64, _c_i64, _c_p, _c_i64, _c_p, _c_i64, _c_int,
         _F3, _F3, _c_f, _c_f, _c_p, _c_sz, _c_p]),
    "totri_backprojection_confocal_velten_workspace": (_c_sz, [_c_i64] * 3),
    "totri_backprojection_confocal_velten_forward": (
        _c_int,
        [_c_p, _c_i64, _c_i64, _c_i64, _c_p, _c_p, _c_i64, _c_p, _c_i64, _c_i64, _c_i64,
         _F3, _F3, _c_f, _c_f, _c_p, _c_sz, _c_p]),
    "totri_backprojection_confocal_velten_noalloc_max_bytes": (_c_sz, []),
    "totri_backprojection_confocal_velten_forward_noalloc": (
        _c_int,
        [_c_p, _c_i64, _c_i64, _c_i64, _c_p, _c_p, _c_i64, _c_p, _c_i64, _c_i64, _c_i64,
         _F3, _F3, _c_f, _c_f, _c_p]),
    "totri_backprojection_exhaustive_velten_workspace": (_c_sz, [_c_i64] * 4),
    "totri_backprojection_exhaustive_velten_forward": (
        _c_int,
        [_c_p, _c_i64, _c_i64, _c_i64, _c_i64, _c_p, _c_p, _c_p, _c_i64, _c_p, _c_i64,
         _c_p, _c_i64, _c_i64, _c_i64, _F3, _F3, _c_f, _c_f, _c_p, _c_sz, _c_p]),
    "totri_backprojection_exhaustive_arellano_workspace": (_c_sz, [_c_i64] * 7),
    "totri_backprojection_exhaustive_arellano_forward": (
        _c_int,
        [_c_p, _c_i64, _c_i64, _c_i64, _c_i64, _c_p, _c_p, _c_p, _c_i64, _c_p, _c_i64,
         _c_p, _c_i64, _c_i64, _c_i64, _F3, _F3, _c_f, _c_f, _c_p, _c_sz, _c_p]),
    "totri_volume_render_exhaustive_forward_workspace": (_c_sz, [_c_i64] * 4),
    "totri_volume_render_exhaustive_forward": (
        _c_int,
        [_c_p, _c_i64, _c_i64, _c_i64, _c_i64, _c_p, _c_i64, _c_p, _c_i64, _c_p, _c_i64,
         _F3, _F3, _c_f, _c_f, _c_p, _c_sz, _c_p]),
    "totri_volume_render_exhaustive_grad_volume_workspace": (_c_sz, [_c_i64] * 4),
    "totri_volume_render_exhaustive_grad_volume": (
        _c_int,
        [_c_p, _c_i64, _c_i64, _c_i64, _c_i64, _c_p, _c_p, _c_p, _c_i64, _c_i64, _c_i64,
         _F3, _F3, _c_f, _c_f, _c_p, _c_sz, _c_p]),
    "totri_rbf2volume_gaussian_forward_workspace": (
        _c_sz, [_c_i64, _c_i64, _c_i64, _c_i64, _c_i64, _c_i64, _c_f]),
    "totri_rbf2volume_gaussian_forward": (
        _c_int,
        [_c_p, _c_i64, _c_i64, _c_i64, _c_p, _c_i64, _c_i64, _c_i64, _F3, _F3, _c_f, _c_p,
         _c_sz, _c_p]),
    "totri_rbf2volume_gaussian_grad": (
        _c_int,
        [_c_p, _c_i64, _c_i64, _c_i64, _c_p, _c_p, _c_p, _c_i64, _c_i64, _c_i64, _F3, _F3,
         _c_f, _c_p]),
    "totri_confocal_job_workspace": (_c_sz, [_c_p]),
    "totri_confocal_batch_table_bytes": (_c_sz, []),
    "totri_confocal_batch_prepare": (_c_int, [_c_p, _c_int, _c_p, _c_sz]),
    "totri_confocal_batch_run": (_c_int, [_c_p, _c_p, _c_sz, _c_p]),
    "totri_fk_prepare": (_c_int, [_c_p, _c_i64, _c_i64, _c_i64, _c_p, _c_int, _c_p]),
    "totri_fk_stolt": (
        _c_int,
        [_c_p, _c_i64, _c_i64, _c_i64, _c_i64, ctypes.c_double, ctypes.c_double,
         ctypes.c_double, _c_int, _c_int, _c_p, _c_p]),
    "totri_fk_magnitude": (_c_int, [_c_p, _c_i64, _c_i64, _c_i64, _c_i64, _c_p, _c_p]),
    "totri_fk_migration_workspace": (_c_sz, [_c_i64] * 3),
    "totri_fk_migration": (
        _c_int,
        [_c_p, _c_i64, _c_i64, _c_i64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
         _c_p, _c_p, _c_sz, _c_p]),
    "totri_debug_mirror_workspace": (_c_sz, [_c_i64]),
    "totri_debug_mirror_info": (_c_int, [_c_p, _c_i64, _c_p, _c_sz, _c_p, _c_p]),
    "totri_debug_a1_walks": (
        _c_int, [_c_i64, _c_i64, _c_i64, _c_i64, _c_i64, _F3, _F3, _c_f, _c_f]),
    "totri_debug_r1_paths": (_c_int, [_c_i64, _c_i64, _c_i64, _c_i64, _c_i64, _F3, _F3]),
    "totri_debug_sqrt_mid_mismatches": (_c_int, [ctypes.c_uint32, ctypes.c_uint32, _c_p, _c_p]),
    "totri_debug_confocal_samples": (
        _c_int,
        [_c_p, _c_i64, _c_p, _c_i64, _c_i64, _c_i64, _c_i64, _c_i64, _F3, _F3,
         _c_f, _c_f, _c_p, _c_p, _c_p, _c_p]),
}

TOTRI_JOB_RENDER = 0
TOTRI_JOB_BACKPROJECT = 1
TOTRI_BATCH_MAX_JOBS = 64


class ConfocalJob(ctypes.Structure):
    """totri_confocal_job of include/totri_hip.h (one request of a batched step)."""
    _fields_ = [
        ("kind", ctypes.c_int),
        ("slice", ctypes.c_int),
        ("volume", _c_p),
        ("measurement_points", _c_p),
        ("transient", _c_p),
        ("X", _c_i64), ("Y", _c_i64), ("Z", _c_i64), ("N", _c_i64), ("T", _c_i64),
        ("volume_origin", _F3),
        ("volume_size", _F3),
        ("bin_width", _c_f),
        ("bin_offset", _c_f),
        ("slices", ctypes.c_int),
        ("slice_run", ctypes.c_int),
        ("workspace", _c_p),
        ("workspace_bytes", _c_sz),
    ]


_lock = threading.Lock()
_lib = None


class TotriLibraryError(RuntimeError):
    """The HIP library is missing or failed to load."""


def load() -> ctypes.CDLL:
    """Load libtotri_hip.so (once) and declare every exported signature."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise TotriLibraryError(
                f"{LIB_PATH} is not built; run `python -m totrilib_amd.build` "
                "(or __graft_entry__.build()) first")
        try:
            lib = ctypes.CDLL(LIB_PATH)
        except OSError as e:
            raise TotriLibraryError(f"cannot load {LIB_PATH}: {e}") from e
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def last_error() -> str:
    msg = load().totri_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str) -> None:
    """Map a status code to the exception the reference would raise."""
    if rc == TOTRI_OK:
        return
    msg = f"{what}: {last_error()}"
    if rc == TOTRI_EINVAL:
        raise ValueError(msg)
    raise RuntimeError(msg)


def f3(v) -> ctypes.Array:
    v = tuple(float(x) for x in v)
    if len(v) != 3:
        raise ValueError(f"expected 3 entries, got {len(v)}")
    return _F3(*v)
