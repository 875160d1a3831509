"""ctypes binding of liboracle.so — TEST INFRASTRUCTURE ONLY.

The oracle is the CPU restatement of mesh_to_sdf's hot path (sdf_oracle.c).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product package never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

OK = 0
EMPTY_RESULT = 1
ERR_NON_FINITE_QUERY = -1
ERR_INDEX_COUNT = -2
ERR_VERTEX_COUNT = -3
ERR_INDEX_OUT_OF_RANGE = -4
ERR_EMPTY_MESH = -5
ERR_NAN_DISTANCE = -6
ERR_INVALID_ARGUMENT = -7
ERR_INVALID_GRID = -10

TOPO_LIST, TOPO_STRIP = 0, 1
SIGN_RAYCAST, SIGN_NORMAL = 0, 1
ACCEL_NONE, ACCEL_BVH, ACCEL_RTREE, ACCEL_RTREE_BVH = 0, 1, 2, 3
GRID_SEQUENTIAL, GRID_THREADED, GRID_FRONTIER, GRID_EXACT = 0, 1, 2, 3
(FAULT_NONE, FAULT_SKIP_DIRECTION, FAULT_DROP_DEFERRED, FAULT_PACKED_ORDER, FAULT_BOUND_MARGIN_SIGN,
 FAULT_NO_SECOND_RELEASE) = 0, 1, 2, 3, 4, 5
BOUND_OFF, BOUND_EXACT_DIRECTION, BOUND_RANDOM_DIRECTION = 0, 1, 2
NO_TRIANGLE = 0xFFFFFFFF

_f = ctypes.POINTER(ctypes.c_float)
_u32 = ctypes.POINTER(ctypes.c_uint32)
_u64 = ctypes.POINTER(ctypes.c_uint64)
_u8 = ctypes.POINTER(ctypes.c_uint8)
_d = ctypes.POINTER(ctypes.c_double)


class GridStats(ctypes.Structure):
    _fields_ = [
        ("seeded_cells", ctypes.c_uint64),
        ("seed_updates", ctypes.c_uint64),
        ("pops", ctypes.c_uint64),
        ("rays", ctypes.c_uint64),
        ("ray_candidates", ctypes.c_uint64),
        ("rounds", ctypes.c_uint64),
        ("evals", ctypes.c_uint64),
    ]


def build() -> str:
    """Compile liboracle.so with its Makefile (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "orc_triangle_bounding_box": (None, [_f, _f, _f, _f, _f]),
            "orc_closest_point_triangle": (None, [_f, _f, _f, _f, _f]),
            "orc_closest_point_segment": (None, [_f, _f, _f, _f]),
            "orc_point_triangle_distance": (ctypes.c_float, [_f, _f, _f, _f]),
            "orc_point_triangle_distance2": (ctypes.c_float, [_f, _f, _f, _f]),
            "orc_point_triangle_signed_distance": (ctypes.c_float, [_f, _f, _f, _f]),
            "orc_ray_triangle_intersection_aligned": (ctypes.c_int, [_f, _f, _f, _f, ctypes.c_int, _f]),
            "orc_ray_triangle_intersection_generic": (ctypes.c_int, [_f, _f, _f, _f, _f, _f]),
            "orc_baseline_point_triangle_distance": (ctypes.c_float, [_f, _f, _f, _f]),
            "orc_approx_eq_f32": (ctypes.c_int, [ctypes.c_float, ctypes.c_float, ctypes.c_int32, ctypes.c_float]),
            "orc_compare_distances": (ctypes.c_int, [ctypes.c_float, ctypes.c_float]),
            "orc_decode_topology": (ctypes.c_int, [ctypes.c_uint64, _u32, ctypes.c_uint64, ctypes.c_int, _u32,
                                                   ctypes.c_uint64, _u64]),
            "orc_grid_from_bounding_box": (None, [_f, _f, _u64, _f, _f]),
            "orc_grid_cell_center": (None, [_f, _f, _u64, _f]),
            "orc_grid_snap": (ctypes.c_int, [_f, _f, _u64, _f, _u64]),
            "orc_generate_sdf": (ctypes.c_int, [_f, ctypes.c_uint64, _u32, ctypes.c_uint64, ctypes.c_int, _f,
                                                ctypes.c_uint64, ctypes.c_int, ctypes.c_int, _f, ctypes.c_int, _u64]),
            "orc_generate_grid_sdf": (ctypes.c_int, [_f, ctypes.c_uint64, _u32, ctypes.c_uint64, ctypes.c_int, _f, _f,
                                                     _u64, ctypes.c_int, _f, ctypes.c_int, ctypes.c_int, _f,
                                                     ctypes.POINTER(GridStats)]),
            "orc_generate_grid_sdf_ex": (ctypes.c_int, [_f, ctypes.c_uint64, _u32, ctypes.c_uint64, ctypes.c_int, _f,
                                                        _f, _u64, ctypes.c_int, _f, _u32, ctypes.c_int, ctypes.c_int,
                                                        _f, ctypes.POINTER(GridStats)]),
            "orc_grid_fixpoint_check": (ctypes.c_int, [_f, ctypes.c_uint64, _u32, ctypes.c_uint64, ctypes.c_int, _f,
                                                       _f, _u64, ctypes.c_int, _f, _u32, ctypes.c_int, _u64]),
            "orc_set_fault": (None, [ctypes.c_int]),
            "orc_set_bound_check": (None, [ctypes.c_int]),
            "orc_bound_check_counts": (None, [_u64]),
            "orc_nearest_bvh": (None, [_f, _u32, ctypes.c_uint64, _f, ctypes.c_uint64, _f, _u32, ctypes.c_int]),
            "orc_query_vector_to_closest_point": (ctypes.c_int, [_f, ctypes.c_uint64, _u32, ctypes.c_uint64,
                                                                 ctypes.c_int, _f, ctypes.c_uint64, _f, _u8,
                                                                 ctypes.c_int]),
            "orc_nearest_bruteforce": (None, [_f, _u32, ctypes.c_uint64, _f, ctypes.c_uint64, _f, _u32, ctypes.c_int]),
            "orc_fp64_distance": (None, [_f, _u32, ctypes.c_uint64, _f, ctypes.c_uint64, _d, ctypes.c_int]),
            "orc_ray_parity_bruteforce": (None, [_f, _u32, ctypes.c_uint64, _f, ctypes.c_uint64, _u8, ctypes.c_int]),
            "orc_check_mesh_triangle_list": (ctypes.c_int, [_f, ctypes.c_uint64, _u32, ctypes.c_uint64]),
            "orc_check_mesh_triangle_list_at": (ctypes.c_int, [_f, ctypes.c_uint64, _u32, ctypes.c_uint64,
                                                                ctypes.POINTER(ctypes.c_uint64)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _fp(a):
    return None if a is None else a.ctypes.data_as(_f)


def _v3(p):
    return np.ascontiguousarray(p, dtype=np.float32).reshape(3)


def _ptr3(p):
    a = _v3(p)
    return a, a.ctypes.data_as(_f)


def threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


# ---------------------------------------------------------------- geometry

def triangle_bounding_box(a, b, c):
    lo = np.zeros(3, np.float32)
    hi = np.zeros(3, np.float32)
    (A, pa), (B, pb), (C, pc) = _ptr3(a), _ptr3(b), _ptr3(c)
    lib().orc_triangle_bounding_box(pa, pb, pc, _fp(lo), _fp(hi))
    return lo, hi


def closest_point_triangle(p, a, b, c):
    out = np.zeros(3, np.float32)
    arrs = [_ptr3(x) for x in (p, a, b, c)]
    lib().orc_closest_point_triangle(*[x[1] for x in arrs], _fp(out))
    return out


def closest_point_segment(p, a, b):
    out = np.zeros(3, np.float32)
    arrs = [_ptr3(x) for x in (p, a, b)]
    lib().orc_closest_point_segment(*[x[1] for x in arrs], _fp(out))
    return out


def point_triangle_distance(p, a, b, c):
    arrs = [_ptr3(x) for x in (p, a, b, c)]
    return np.float32(lib().orc_point_triangle_distance(*[x[1] for x in arrs]))


def point_triangle_distance2(p, a, b, c):
    arrs = [_ptr3(x) for x in (p, a, b, c)]
    return np.float32(lib().orc_point_triangle_distance2(*[x[1] for x in arrs]))


def point_triangle_signed_distance(p, a, b, c):
    arrs = [_ptr3(x) for x in (p, a, b, c)]
    return np.float32(lib().orc_point_triangle_signed_distance(*[x[1] for x in arrs]))


def ray_triangle_intersection_aligned(origin, a, b, c, axis):
    arrs = [_ptr3(x) for x in (origin, a, b, c)]
    t = ctypes.c_float(0)
    hit = lib().orc_ray_triangle_intersection_aligned(*[x[1] for x in arrs], int(axis), ctypes.byref(t))
    return np.float32(t.value) if hit else None


def ray_triangle_intersection_generic(origin, direction, a, b, c):
    arrs = [_ptr3(x) for x in (origin, direction, a, b, c)]
    t = ctypes.c_float(0)
    hit = lib().orc_ray_triangle_intersection_generic(*[x[1] for x in arrs], ctypes.byref(t))
    return np.float32(t.value) if hit else None


def baseline_point_triangle_distance(p, a, b, c):
    arrs = [_ptr3(x) for x in (p, a, b, c)]
    return np.float32(lib().orc_baseline_point_triangle_distance(*[x[1] for x in arrs]))


def approx_eq_f32(a, b, ulps=2, epsilon=1e-6):
    return bool(lib().orc_approx_eq_f32(float(np.float32(a)), float(np.float32(b)), int(ulps),
                                        float(np.float32(epsilon))))


def compare_distances(a, b):
    """-1 Less, 0 Equal, 1 Greater; raises where the reference panics."""
    r = lib().orc_compare_distances(float(np.float32(a)), float(np.float32(b)))
    if r == 2:
        raise ValueError("NaN distance")
    return r


# ---------------------------------------------------------------- topology

def decode_topology(n_vertices, indices, topology):
    idx = None if indices is None else np.ascontiguousarray(indices, dtype=np.uint32)
    ni = 0 if idx is None else idx.size
    nt = ctypes.c_uint64(0)
    pidx = None if idx is None else idx.ctypes.data_as(_u32)
    s = lib().orc_decode_topology(n_vertices, pidx, ni, topology, None, 0, ctypes.byref(nt))
    if s:
        return s, None
    tris = np.zeros((nt.value, 3), np.uint32)
    s = lib().orc_decode_topology(n_vertices, pidx, ni, topology, tris.ctypes.data_as(_u32), nt.value,
                                  ctypes.byref(nt))
    return s, (tris if s == 0 else None)


# ---------------------------------------------------------------- grid

def grid_from_bounding_box(bmin, bmax, count):
    first = np.zeros(3, np.float32)
    size = np.zeros(3, np.float32)
    cnt = np.ascontiguousarray(count, dtype=np.uint64)
    (A, pa), (B, pb) = _ptr3(bmin), _ptr3(bmax)
    lib().orc_grid_from_bounding_box(pa, pb, cnt.ctypes.data_as(_u64), _fp(first), _fp(size))
    return first, size


def grid_cell_center(first, size, cell):
    out = np.zeros(3, np.float32)
    c = np.ascontiguousarray(cell, dtype=np.uint64)
    (A, pa), (B, pb) = _ptr3(first), _ptr3(size)
    lib().orc_grid_cell_center(pa, pb, c.ctypes.data_as(_u64), _fp(out))
    return out


def grid_snap(first, size, count, p):
    cell = np.zeros(3, np.uint64)
    cnt = np.ascontiguousarray(count, dtype=np.uint64)
    (A, pa), (B, pb), (P, pp) = _ptr3(first), _ptr3(size), _ptr3(p)
    r = lib().orc_grid_snap(pa, pb, cnt.ctypes.data_as(_u64), pp, cell.ctypes.data_as(_u64))
    return r, cell


# ---------------------------------------------------------------- entry points

def _mesh_args(vertices, indices):
    v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
    idx = None if indices is None else np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1)
    return v, idx


def generate_sdf(vertices, indices, topology, queries, accel, sign=SIGN_RAYCAST, n_threads=None):
    """Returns (status, distances)."""
    v, idx = _mesh_args(vertices, indices)
    q = np.ascontiguousarray(queries, dtype=np.float32).reshape(-1, 3)
    out = np.zeros(len(q), np.float32)
    n = ctypes.c_uint64(0)
    s = lib().orc_generate_sdf(_fp(v), len(v), None if idx is None else idx.ctypes.data_as(_u32),
                               0 if idx is None else idx.size, topology, _fp(q), len(q), accel, sign, _fp(out),
                               n_threads or threads(), ctypes.byref(n))
    return s, out[: n.value]


def generate_grid_sdf(vertices, indices, topology, first_cell, cell_size, cell_count, sign=SIGN_RAYCAST,
                      n_heaps=1, schedule=GRID_SEQUENTIAL, want_seed=False, want_tri=False):
    """Returns (status, distances, stats[, seeded preheap distances][, per-cell triangle])."""
    v, idx = _mesh_args(vertices, indices)
    cnt = np.ascontiguousarray(cell_count, dtype=np.uint64)
    ncells = int(cnt[0]) * int(cnt[1]) * int(cnt[2])
    out = np.zeros(ncells, np.float32)
    seed = np.zeros(ncells, np.float32) if want_seed else None
    tri = np.zeros(ncells, np.uint32) if want_tri else None
    st = GridStats()
    (F, pf), (S, ps) = _ptr3(first_cell), _ptr3(cell_size)
    s = lib().orc_generate_grid_sdf_ex(_fp(v), len(v), None if idx is None else idx.ctypes.data_as(_u32),
                                       0 if idx is None else idx.size, topology, pf, ps, cnt.ctypes.data_as(_u64),
                                       sign, _fp(out), None if tri is None else tri.ctypes.data_as(_u32), n_heaps,
                                       schedule, _fp(seed), ctypes.byref(st))
    res = [s, out, st]
    if want_seed:
        res.append(seed)
    if want_tri:
        res.append(tri)
    return tuple(res)


def grid_fixpoint_check(vertices, indices, topology, first_cell, cell_size, cell_count, sign, values, tri_of,
                        n_threads=None):
    """Violations of the propagation's fixpoint (orc_grid_fixpoint_check):
    dict with value_mismatch, not_fixpoint, no_triangle, no_triangle_value."""
    v, idx = _mesh_args(vertices, indices)
    cnt = np.ascontiguousarray(cell_count, dtype=np.uint64)
    vals = np.ascontiguousarray(values, dtype=np.float32).reshape(-1)
    tri = np.ascontiguousarray(tri_of, dtype=np.uint32).reshape(-1)
    assert vals.size == tri.size == int(cnt[0]) * int(cnt[1]) * int(cnt[2])
    counts = np.zeros(4, np.uint64)
    (F, pf), (S, ps) = _ptr3(first_cell), _ptr3(cell_size)
    s = lib().orc_grid_fixpoint_check(_fp(v), len(v), None if idx is None else idx.ctypes.data_as(_u32),
                                      0 if idx is None else idx.size, topology, pf, ps, cnt.ctypes.data_as(_u64),
                                      sign, _fp(vals), tri.ctypes.data_as(_u32), n_threads or threads(),
                                      counts.ctypes.data_as(_u64))
    assert s == OK, s
    return dict(zip(("value_mismatch", "not_fixpoint", "no_triangle", "no_triangle_value"), map(int, counts)))


def set_fault(kind):
    """Test-only: inject a fault into the GRID_FRONTIER schedule (FAULT_*)."""
    lib().orc_set_fault(kind)


def set_bound_check(mode):
    """Test-only: check the device's supporting-plane skip on every
    GRID_FRONTIER pair (orc_set_bound_check; BOUND_*); resets the counts."""
    lib().orc_set_bound_check(mode)


def bound_check_counts():
    """Pairs checked, pairs the bound skips, skipped pairs that would improve."""
    c = np.zeros(3, np.uint64)
    lib().orc_bound_check_counts(c.ctypes.data_as(_u64))
    return dict(zip(("checked", "skipped", "wrong"), map(int, c)))


def query_vector_to_closest_point(vertices, indices, topology, queries, n_threads=None):
    v, idx = _mesh_args(vertices, indices)
    q = np.ascontiguousarray(queries, dtype=np.float32).reshape(-1, 3)
    vec = np.zeros((len(q), 3), np.float32)
    inside = np.zeros(len(q), np.uint8)
    s = lib().orc_query_vector_to_closest_point(_fp(v), len(v), None if idx is None else idx.ctypes.data_as(_u32),
                                                0 if idx is None else idx.size, topology, _fp(q), len(q), _fp(vec),
                                                inside.ctypes.data_as(_u8), n_threads or threads())
    return s, vec, inside.astype(bool)


def nearest_bruteforce(vertices, tris, points, n_threads=None):
    v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(tris, dtype=np.uint32).reshape(-1, 3)
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    d2 = np.zeros(len(p), np.float32)
    ti = np.zeros(len(p), np.uint32)
    lib().orc_nearest_bruteforce(_fp(v), t.ctypes.data_as(_u32), len(t), _fp(p), len(p), _fp(d2),
                                 ti.ctypes.data_as(_u32), n_threads or threads())
    return d2, ti


def nearest_bvh(vertices, tris, points, n_threads=None):
    """nearest_bruteforce's (d2, triangle) through the oracle's BVH."""
    v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(tris, dtype=np.uint32).reshape(-1, 3)
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    d2 = np.zeros(len(p), np.float32)
    ti = np.zeros(len(p), np.uint32)
    lib().orc_nearest_bvh(_fp(v), t.ctypes.data_as(_u32), len(t), _fp(p), len(p), _fp(d2),
                          ti.ctypes.data_as(_u32), n_threads or threads())
    return d2, ti


def grid_centres(first_cell, cell_size, cell_count):
    """Cell centres first + i * size in FP32 with no FMA (grid.rs:289-295), z fastest."""
    f = _v3(first_cell)
    sz = _v3(cell_size)
    n = [int(c) for c in cell_count]
    ax = [f[k] + np.arange(n[k], dtype=np.float32) * sz[k] for k in range(3)]
    X, Y, Z = np.meshgrid(ax[0], ax[1], ax[2], indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1).astype(np.float32)


def fp64_distance(vertices, tris, points, n_threads=None):
    v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(tris, dtype=np.uint32).reshape(-1, 3)
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    out = np.zeros(len(p), np.float64)
    lib().orc_fp64_distance(_fp(v), t.ctypes.data_as(_u32), len(t), _fp(p), len(p), out.ctypes.data_as(_d),
                            n_threads or threads())
    return out


def ray_parity_bruteforce(vertices, tris, points, n_threads=None):
    v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(tris, dtype=np.uint32).reshape(-1, 3)
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    out = np.zeros((len(p), 3), np.uint8)
    lib().orc_ray_parity_bruteforce(_fp(v), t.ctypes.data_as(_u32), len(t), _fp(p), len(p),
                                    out.ctypes.data_as(_u8), n_threads or threads())
    return out


def check_mesh_triangle_list(vertices, faces):
    v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
    f = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1)
    return lib().orc_check_mesh_triangle_list(_fp(v), len(v), f.ctypes.data_as(_u32), f.size)


def check_mesh_triangle_list_at(vertices, faces):
    """(kind, where): kind 0..5 as check_mesh_triangle_list; where is the failing
    vertex index (2), face index (3) or edge index 3 * face + k (4, 5)."""
    v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
    f = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1)
    w = ctypes.c_uint64(0)
    k = lib().orc_check_mesh_triangle_list_at(_fp(v), len(v), f.ctypes.data_as(_u32), f.size, ctypes.byref(w))
    return k, int(w.value)
