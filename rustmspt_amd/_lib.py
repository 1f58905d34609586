"""ctypes binding of libs2hip.so (the C ABI declared in include/s2hip.h).

The library is loaded from this package directory.  There is no fallback: if the
shared object is missing or a call fails, an S2HipError is raised, so the HIP
path is the only thing that can run.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libs2hip.so"

S2H_OK = 0
STATUS_NAMES = {1: "ERR_ARG", 2: "ERR_HIP", 3: "ERR_BUDGET", 4: "ERR_SIZE", 5: "ERR_UNSUPPORTED", 6: "ERR_IO"}

_u64p = C.POINTER(C.c_uint64)
_i64p = C.POINTER(C.c_int64)
_u32p = C.POINTER(C.c_uint32)
_u8p = C.POINTER(C.c_uint8)
_f64p = C.POINTER(C.c_double)
_f32p = C.POINTER(C.c_float)
_vp = C.c_void_p

# name -> (restype, argtypes); every symbol include/s2hip.h declares
SIGNATURES = {
    "s2h_create": (C.c_int32, [C.c_int, C.c_uint64, C.POINTER(_vp)]),
    "s2h_destroy": (None, [_vp]),
    "s2h_last_error": (C.c_char_p, []),
    "s2h_device_name": (C.c_int32, [_vp, C.c_char_p, C.c_uint64]),
    "s2h_bytes_in_use": (C.c_int32, [_vp, _u64p]),
    "s2h_profile": (C.c_int32, [_vp, C.c_int]),
    "s2h_profile_read": (C.c_int32, [_vp, _f64p, C.c_uint64]),
    "s2h_valu_probe": (C.c_int32, [_vp, _f64p]),
    "s2h_mesh_mc_counts": (C.c_int32, [_vp, _f64p, C.c_uint64, _u32p, C.c_uint64, _f64p, _f64p, C.c_uint32,
                                       C.c_uint64, C.c_uint64, _u64p, C.c_uint64]),
    "s2h_mesh_mc_s2": (C.c_int32, [_vp, _f64p, C.c_uint64, _u32p, C.c_uint64, _f64p, _f64p, C.c_uint32,
                                   C.c_uint64, C.c_uint64, C.c_double, _f64p, C.c_uint64]),
    "s2h_mesh_from_stl": (C.c_int32, [_vp, C.POINTER(_vp), _u64p, C.c_uint32, C.POINTER(_vp)]),
    "s2h_mesh_load_stl": (C.c_int32, [_vp, C.POINTER(C.c_char_p), C.c_uint32, C.POINTER(_vp)]),
    "s2h_mesh_from_arrays": (C.c_int32, [_vp, _f64p, C.c_uint64, _u32p, C.c_uint64, C.POINTER(_vp)]),
    "s2h_mesh_info": (C.c_int32, [_vp, _u64p, _u64p, _u32p, _f64p, _f64p]),
    "s2h_mesh_read": (C.c_int32, [_vp, _vp, _f64p, C.c_uint64, _u32p, C.c_uint64, _u32p, C.c_uint64]),
    "s2h_mesh_voxelize": (C.c_int32, [_vp, _vp, _f64p, C.c_double, _u64p, C.POINTER(_vp), _u64p]),
    "s2h_mesh_free": (None, [_vp, _vp]),
    "s2h_mesh_set_vertices": (C.c_int32, [_vp, _vp, C.c_uint64, _f64p, C.c_uint64]),
    "s2h_mesh_mc_counts_resident": (C.c_int32, [_vp, _vp, _f64p, _f64p, C.c_uint32, C.c_uint64, C.c_uint64, _u64p,
                                                C.c_uint64]),
    "s2h_ordered_sum": (C.c_int32, [_vp, _f64p, C.c_uint64, _u64p, C.c_uint32, _f64p]),
    "s2h_volume_probe": (C.c_int32, [_vp, _vp]),
    "s2h_volume_read": (C.c_int32, [_vp, _vp, C.c_uint64, _vp]),
    "s2h_vol_from_samples": (C.c_int32, [_vp, _vp, C.c_uint64, _vp, _vp, C.POINTER(_vp), _i64p, _u64p]),
    "s2h_vol_load": (C.c_int32, [_vp, _vp, _vp, C.POINTER(_vp), _vp, _i64p, _u64p]),
    "s2h_vol_from_packed": (C.c_int32, [_vp, _u64p, C.c_uint64, _u64p, C.POINTER(_vp)]),
    "s2h_vol_from_bytes": (C.c_int32, [_vp, _u8p, _u64p, C.POINTER(_vp)]),
    "s2h_vol_read_packed": (C.c_int32, [_vp, _vp, _u64p, C.c_uint64]),
    "s2h_vol_dims": (C.c_int32, [_vp, _u64p]),
    "s2h_vol_occupied": (C.c_int32, [_vp, _vp, _u64p]),
    "s2h_vol_free": (None, [_vp, _vp]),
    "s2h_voxelize": (C.c_int32, [_vp, _f64p, C.c_uint64, _u32p, C.c_uint64, _u32p, C.c_uint32, _f64p,
                                 C.c_double, _u64p, C.POINTER(_vp), _u64p]),
    "s2h_shell_count": (C.c_int32, [_u64p, C.c_uint32, C.c_double, _u64p]),
    "s2h_shell_offsets": (C.c_int32, [_u64p, C.c_uint32, C.c_double, _i64p, C.c_uint64, _u64p, C.c_uint64]),
    "s2h_exact_counts": (C.c_int32, [_vp, _vp, C.c_uint32, C.c_double, _u64p, C.c_uint64]),
    "s2h_exact_s2": (C.c_int32, [_vp, _vp, C.c_uint32, C.c_double, _f64p, C.c_uint64]),
    "s2h_exact_tune_info": (C.c_int32, [_vp, _u64p, C.c_uint32, C.c_double, _i64p, C.c_uint64]),
    "s2h_exact_tune_set": (C.c_int32, [_vp, _u64p, C.c_uint32, C.c_double, C.c_int32, C.c_int32]),
    "s2h_set_autotune": (C.c_int32, [_vp, C.c_int32]),
    "s2h_exact_method_set": (C.c_int32, [_vp, C.c_int32]),
    "s2h_fft_padded_dims": (C.c_int32, [_u64p, C.c_uint32, C.c_double, _u64p]),
    "s2h_fft32_counts": (C.c_int32, [_vp, _vp, C.c_uint32, C.c_double, _f32p, C.c_uint64]),
    "s2h_fft32_s2": (C.c_int32, [_vp, _vp, C.c_uint32, C.c_double, _f64p, C.c_uint64]),
    "s2h_mc_counts": (C.c_int32, [_vp, _vp, C.c_uint32, C.c_double, C.c_uint64, C.c_uint64, _u64p, C.c_uint64]),
    "s2h_mc_counts_noalloc": (C.c_int32, [_vp, _vp, C.c_uint32, C.c_double, C.c_uint64, C.c_uint64, _u64p,
                                          C.c_uint64]),
    "s2h_mc_s2": (C.c_int32, [_vp, _vp, C.c_uint32, C.c_double, C.c_uint64, C.c_uint64, _f64p, C.c_uint64]),
    "s2h_fill_missing": (C.c_int32, [_f64p, _u8p, C.c_uint64, C.c_double]),
    "s2h_serve_create": (C.c_int32, [C.c_int, C.c_uint64, C.c_uint32, C.POINTER(_vp)]),
    "s2h_serve_destroy": (None, [_vp]),
    "s2h_serve_submit": (C.c_int32, [_vp, _vp, _u64p]),
    "s2h_serve_poll": (C.c_int32, [_vp, C.c_uint64, C.c_int32, _vp, _f64p, C.c_uint64]),
    "s2h_serve_stats": (C.c_int32, [_vp, _u64p, C.c_uint64]),
    "s2h_seed_create": (C.c_int32, [_vp, _f64p, C.c_uint64, _u32p, C.c_uint64, _u32p, C.c_uint32, _u32p, _u32p,
                                    C.c_uint32, _f64p, C.c_double, _u64p, C.c_uint32, C.POINTER(_vp)]),
    "s2h_seed_accept": (C.c_int32, [_vp, _vp, C.c_uint32, _f64p, _f64p, C.c_uint64]),
    "s2h_seed_free": (None, [_vp, _vp]),
    "s2h_seed_volume": (C.c_int32, [_vp, C.POINTER(_vp)]),
    "s2h_batch_eval": (C.c_int32, [_vp, C.POINTER(_vp), C.c_uint32, _vp, C.c_uint32, _f64p, C.c_uint64]),
    "s2h_stl_load": (C.c_int32, [C.c_char_p, C.POINTER(_f64p), _u64p, C.POINTER(_u32p), _u64p]),
    "s2h_free": (None, [_vp]),
    "s2h_mesh_granules": (C.c_int32, [_u32p, C.c_uint64, C.c_uint64, _u32p, _u32p]),
    "s2h_mesh_volume_in_bbox": (C.c_int32, [_f64p, C.c_uint64, _u32p, C.c_uint64, _f64p, _f64p, _f64p]),
    "s2h_volume_fraction_in_bbox": (C.c_int32, [_f64p, C.c_uint64, _u32p, C.c_uint64, _f64p, _f64p, _f64p]),
}


class Candidate(C.Structure):
    """s2h_candidate (include/s2hip.h)."""
    _fields_ = [("seed", C.c_uint32), ("particle", C.c_uint32), ("new_xyz", _f64p)]


class Request(C.Structure):
    """s2h_request (include/s2hip.h)."""
    _fields_ = [("words", _u64p), ("nwords", C.c_uint64), ("dims", C.c_uint64 * 3), ("r_max", C.c_uint32),
                ("method", C.c_int32), ("pitch", C.c_double), ("tag", C.c_uint64)]


class Result(C.Structure):
    """s2h_result (include/s2hip.h)."""
    _fields_ = [("ready", C.c_int32), ("status", C.c_int32), ("deferrals", C.c_uint32), ("reserved", C.c_uint32),
                ("tag", C.c_uint64), ("footprint", C.c_uint64), ("t_submit", C.c_double), ("t_start", C.c_double),
                ("t_done", C.c_double)]


class VolumeSource(C.Structure):
    _fields_ = [("path", C.c_char_p), ("kind", C.c_int32), ("reserved", C.c_int32), ("slice_start", C.c_int64),
                ("slice_end", C.c_int64), ("width", C.c_uint64), ("height", C.c_uint64), ("bits", C.c_uint32),
                ("is_signed", C.c_int32), ("big_endian", C.c_int32)]


class VolumeInfo(C.Structure):
    _fields_ = [("width", C.c_uint64), ("height", C.c_uint64), ("depth", C.c_uint64), ("bits", C.c_uint32),
                ("is_signed", C.c_int32), ("big_endian", C.c_int32), ("reserved", C.c_int32)]


class Phase(C.Structure):
    _fields_ = [("mode", C.c_int32), ("reserved", C.c_int32), ("lo", C.c_int64), ("hi", C.c_int64),
                ("background", C.c_int64)]


class S2HipError(RuntimeError):
    """A failing C-ABI call: status code plus the library's message."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status
        self.message = message


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise S2HipError(-1, f"{LIB_PATH} is not built; run __graft_entry__.build() or "
                                 f"python rustmspt_amd/build.py")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status: int) -> None:
    if status != S2H_OK:
        msg = lib().s2h_last_error()
        raise S2HipError(status, msg.decode() if msg else "")


def ptr(a, t):
    return a.ctypes.data_as(t)
