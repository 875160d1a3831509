"""Loader and ctypes signatures for libsdfx.so (the HIP library built from csrc/).

The library is required: if it is missing or fails to load, every entry point
raises. There is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsdfx.so")
CSRC = os.path.join(HERE, "csrc")

OK = 0
EMPTY_RESULT = 1
MESH_INVALID = 2
ERR_NON_FINITE_QUERY = -1
ERR_INDEX_COUNT = -2
ERR_VERTEX_COUNT = -3
ERR_INDEX_OUT_OF_RANGE = -4
ERR_EMPTY_MESH = -5
ERR_NAN_DISTANCE = -6
ERR_INVALID_ARGUMENT = -7
ERR_DEVICE = -8
ERR_OUT_OF_MEMORY = -9
ERR_INVALID_GRID = -10
ERR_OVER_BUDGET = -11
ERR_CANCELLED = -12

_f = ctypes.POINTER(ctypes.c_float)
_u32 = ctypes.POINTER(ctypes.c_uint32)
_u64 = ctypes.POINTER(ctypes.c_uint64)
_u8 = ctypes.POINTER(ctypes.c_uint8)
_vp = ctypes.c_void_p
_i = ctypes.c_int
_u64v = ctypes.c_uint64

class GridRequest(ctypes.Structure):
    """sdfx_grid_request (include/sdfx.h)."""
    _fields_ = [("vertices", ctypes.c_void_p), ("vertex_count", ctypes.c_uint64), ("indices", ctypes.c_void_p),
                ("index_count", ctypes.c_uint64), ("topology", ctypes.c_int), ("first_cell", ctypes.c_float * 3),
                ("cell_size", ctypes.c_float * 3), ("cell_count", ctypes.c_uint64 * 3), ("out", ctypes.c_void_p),
                ("status", ctypes.c_int)]


class ServerMesh(ctypes.Structure):
    """sdfx_server_mesh (include/sdfx.h)."""
    _fields_ = [("vertices", ctypes.c_void_p), ("vertex_count", ctypes.c_uint64), ("indices", ctypes.c_void_p),
                ("index_count", ctypes.c_uint64), ("topology", ctypes.c_int), ("id", ctypes.c_uint32),
                ("status", ctypes.c_int)]


class PointRequest(ctypes.Structure):
    """sdfx_point_request (include/sdfx.h)."""
    _fields_ = [("mesh", ctypes.c_void_p), ("query_points", ctypes.c_void_p), ("query_count", ctypes.c_uint64),
                ("out", ctypes.c_void_p), ("status", ctypes.c_int)]


class ServerStats(ctypes.Structure):
    """sdfx_server_stats (include/sdfx.h)."""
    _fields_ = [(n, ctypes.c_uint64) for n in ("requests_done", "batches", "max_batch_requests", "resident_bytes",
                                                "inflight_bytes", "peak_inflight_bytes", "meshes", "queued")]


# name -> (restype, argtypes); must match include/sdfx.h
SIGNATURES = {
    "sdfx_last_error": (ctypes.c_char_p, []),
    "sdfx_version": (_i, []),
    "sdfx_generate_sdf": (_i, [_f, _u64v, _u32, _u64v, _i, _f, _u64v, _i, _i, _f, _u64]),
    "sdfx_generate_sdf_device": (_i, [_vp, _u64v, _vp, _u64v, _i, _vp, _u64v, _i, _i, _vp, _u64, _vp]),
    "sdfx_generate_grid_sdf": (_i, [_f, _u64v, _u32, _u64v, _i, _f, _f, _u64, _i, _f]),
    "sdfx_generate_grid_sdf_device": (_i, [_vp, _u64v, _vp, _u64v, _i, _f, _f, _u64, _i, _vp, _vp]),
    "sdfx_generate_grid_sdf_exact": (_i, [_f, _u64v, _u32, _u64v, _i, _f, _f, _u64, _i, _f]),
    "sdfx_generate_grid_sdf_exact_device": (_i, [_vp, _u64v, _vp, _u64v, _i, _f, _f, _u64, _i, _vp, _vp]),
    "sdfx_build_sdf_acceleration": (_i, [_f, _u64v, _u32, _u64v, _i, _i, _i, ctypes.POINTER(_vp)]),
    "sdfx_query_sdf": (_i, [_vp, _f, _u64v, _f, _u64]),
    "sdfx_query_sdf_noalloc": (_i, [_vp, _vp, _u64v, _vp, _u64, _vp]),
    "sdfx_query_sdf_batch_device": (_i, [ctypes.POINTER(PointRequest), _u64v, _vp]),
    "sdfx_query_vector_to_closest_point": (_i, [_vp, _f, _u64v, _f, _u8]),
    "sdfx_free_sdf_acceleration": (None, [_vp]),
    "sdfx_generate_grid_sdf_batch": (_i, [ctypes.POINTER(GridRequest), _u64v, _i]),
    "sdfx_generate_grid_sdf_batch_device": (_i, [ctypes.POINTER(GridRequest), _u64v, _i, _vp]),
    "sdfx_check_mesh_triangle_list": (_i, [_f, _u64v, _u32, _u64v, ctypes.POINTER(_i), _u64, ctypes.c_char_p,
                                          _u64v]),
    "sdfx_server_create": (_i, [_u64v, ctypes.c_uint32, ctypes.POINTER(_vp)]),
    "sdfx_server_destroy": (None, [_vp]),
    "sdfx_server_add_mesh": (_i, [_vp, _vp, _u64v, _vp, _u64v, _i, _u32]),
    "sdfx_server_add_meshes": (_i, [_vp, ctypes.POINTER(ServerMesh), _u64v]),
    "sdfx_server_remove_mesh": (_i, [_vp, ctypes.c_uint32]),
    "sdfx_server_submit_points": (_i, [_vp, ctypes.c_uint32, _vp, _u64v, _i, _i, _vp, _u64]),
    "sdfx_server_submit_grid": (_i, [_vp, ctypes.c_uint32, _f, _f, _u64, _i, _vp, _u64]),
    "sdfx_server_wait": (_i, [_vp, _u64v, ctypes.POINTER(ctypes.c_double)]),
    "sdfx_server_cancel": (_i, [_vp, _u64v, ctypes.POINTER(_i)]),
    "sdfx_server_poll": (_i, [_vp, _u64v, ctypes.POINTER(_i)]),
    "sdfx_server_get_stats": (_i, [_vp, ctypes.POINTER(ServerStats)]),
    "sdfx_profile_enable": (None, [_i]),
    "sdfx_profile_read": (_i, [_i, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), _u64]),
    "sdfx_profile_counter": (ctypes.c_uint64, [ctypes.c_char_p]),
    "sdfx_profile_select": (None, [ctypes.c_char_p, _i]),
    "sdfx_set_tuning": (_i, [ctypes.c_char_p, ctypes.c_int64]),
    "sdfx_get_tuning": (ctypes.c_int64, [ctypes.c_char_p]),
}

_lib = None


class SdfxError(RuntimeError):
    """A status the reference reports by panicking."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[{status}] {message}")
        self.status = status
        self.message = message


def build(jobs: int = 8) -> str:
    """Compile libsdfx.so for gfx950 with hipcc (no GPU needed)."""
    subprocess.run(["make", "-s", "-C", CSRC, f"-j{jobs}"], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C {CSRC}")
        # One HIP runtime per process. torch ships its own libamdhip64 (soname
        # libamdhip64.so.7) and loads it by another file name, so a process that
        # loaded the system runtime first would end up with two, and the first
        # one loses the GPU. Importing torch first makes libsdfx bind to torch's
        # copy through the shared soname.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
        apply_tuning_file()
    return _lib


TUNING_PATH = os.path.join(HERE, "tuning.json")


def apply_tuning_file(path: str = None) -> dict:
    """Applies the knobs of a tuning file written by mesh_to_sdf_amd.autotune
    (default: SDFX_TUNING, else mesh_to_sdf_amd/tuning.json; SDFX_TUNING=none
    skips it). The knobs change launch shapes only, never a result."""
    import json
    path = path or os.environ.get("SDFX_TUNING", TUNING_PATH)
    if path == "none" or not os.path.exists(path):
        return {}
    with open(path) as fh:
        knobs = json.load(fh).get("knobs", {})
    for key, value in knobs.items():
        set_tuning(key, int(value))
    return knobs


def profile_enable(on: bool = True) -> None:
    lib().sdfx_profile_enable(1 if on else 0)


def profile_read(max_entries: int = 128) -> dict:
    """{kernel name: (total ms, launches)} since the last profile_enable()."""
    names = ctypes.create_string_buffer(64 * max_entries)
    ms = (ctypes.c_double * max_entries)()
    cnt = (ctypes.c_uint64 * max_entries)()
    n = lib().sdfx_profile_read(max_entries, names, ms, cnt)
    out = {}
    for i in range(min(n, max_entries)):
        name = names.raw[64 * i:64 * (i + 1)].split(b"\0", 1)[0].decode()
        out[name] = (ms[i], int(cnt[i]))
    return out


def profile_select(kernel_prefix: str = "", counters: bool = True) -> None:
    lib().sdfx_profile_select(kernel_prefix.encode(), 1 if counters else 0)


def profile_counter(name: str) -> int:
    return int(lib().sdfx_profile_counter(name.encode()))


def set_tuning(key: str, value: int) -> None:
    if lib().sdfx_set_tuning(key.encode(), int(value)) != 0:
        raise KeyError(key)


def get_tuning(key: str) -> int:
    return int(lib().sdfx_get_tuning(key.encode()))


class tuned:
    """Context manager setting a tuning knob and restoring it on exit."""

    def __init__(self, key: str, value: int):
        self.key, self.value = key, value

    def __enter__(self):
        self.old = get_tuning(self.key)
        set_tuning(self.key, self.value)

    def __exit__(self, *a):
        set_tuning(self.key, self.old)


def check(status: int) -> int:
    if status < 0:
        msg = lib().sdfx_last_error()
        raise SdfxError(status, msg.decode() if msg else "")
    return status
