"""mesh_to_sdf_amd — mesh_to_sdf's hot path on MI355X (gfx950), Python face.

Mirrors the reference crate's public API (mesh_to_sdf/src/lib.rs:148-485,
grid.rs, generate/grid.rs:265) over the C ABI of libsdfx.so:

    generate_sdf(vertices, Topology.TriangleList(indices), query_points,
                 AccelerationMethod.RtreeBvh)
    generate_grid_sdf(vertices, topology, Grid.from_bounding_box(lo, hi, n),
                      SignMethod.Raycast)
    build_sdf_acceleration / query_sdf / query_vector_to_closest_point_rtree_bvh
    check_mesh_triangle_list(vertices, faces)
    serde.save_to_file / serde.read_from_file (mesh_to_sdf_amd.serde)

Vertices and query points are (N, 3) float32 arrays, indices u16/u32 arrays.
Where the reference panics, an SdfxError carrying the status is raised.
All computation runs in libsdfx.so's HIP kernels; nothing here computes
distances.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import SdfxError

__all__ = [
    "SignMethod", "AccelerationMethod", "ALL_SDF_ACCELERATION_METHODS", "Topology", "Grid", "SnapResult",
    "generate_sdf", "generate_grid_sdf", "build_sdf_acceleration", "query_sdf",
    "query_vector_to_closest_point_rtree_bvh", "SdfAccelerationMesh", "SdfxError", "check_mesh_triangle_list",
    "generate_grid_sdf_batch", "query_sdf_batch", "generate_grid_sdf_exact",
]

_F = ctypes.POINTER(ctypes.c_float)
_U32 = ctypes.POINTER(ctypes.c_uint32)
_U64 = ctypes.POINTER(ctypes.c_uint64)
_U8 = ctypes.POINTER(ctypes.c_uint8)


class SignMethod:
    """lib.rs:252-263 (Raycast is the default)."""
    Raycast = 0
    Normal = 1


@dataclass(frozen=True)
class _Accel:
    kind: int
    sign: int

    def __repr__(self):
        name = ["None", "Bvh", "Rtree", "RtreeBvh"][self.kind]
        if self.kind in (0, 1):
            return f"AccelerationMethod.{name}({'Raycast' if self.sign == 0 else 'Normal'})"
        return f"AccelerationMethod.{name}"


class AccelerationMethod:
    """lib.rs:272-286. `None` is a Python keyword, so the first variant is
    spelled `AccelerationMethod.None_(sign)` (also reachable as
    getattr(AccelerationMethod, "None"))."""

    @staticmethod
    def None_(sign: int = SignMethod.Raycast) -> _Accel:
        return _Accel(0, sign)

    @staticmethod
    def Bvh(sign: int = SignMethod.Raycast) -> _Accel:
        return _Accel(1, sign)

    Rtree = _Accel(2, SignMethod.Normal)
    RtreeBvh = _Accel(3, SignMethod.Raycast)
    default = RtreeBvh


setattr(AccelerationMethod, "None", AccelerationMethod.None_)

# lib.rs:289-296
ALL_SDF_ACCELERATION_METHODS = (
    AccelerationMethod.None_(SignMethod.Raycast),
    AccelerationMethod.None_(SignMethod.Normal),
    AccelerationMethod.Bvh(SignMethod.Raycast),
    AccelerationMethod.Bvh(SignMethod.Normal),
    AccelerationMethod.Rtree,
    AccelerationMethod.RtreeBvh,
)


@dataclass(frozen=True)
class Topology:
    """lib.rs:156-171: TriangleList / TriangleStrip, indices optional."""
    kind: int
    indices: Optional[np.ndarray]

    @staticmethod
    def TriangleList(indices=None) -> "Topology":
        return Topology(0, None if indices is None else np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1))

    @staticmethod
    def TriangleStrip(indices=None) -> "Topology":
        return Topology(1, None if indices is None else np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1))


class SnapResult:
    """grid.rs:10-17."""

    def __init__(self, inside: bool, cell):
        self.inside = inside
        self.cell = tuple(int(c) for c in cell)

    @staticmethod
    def Inside(cell):
        return SnapResult(True, cell)

    @staticmethod
    def Outside(cell):
        return SnapResult(False, cell)

    def __eq__(self, other):
        return isinstance(other, SnapResult) and self.inside == other.inside and self.cell == other.cell

    def __repr__(self):
        return f"SnapResult.{'Inside' if self.inside else 'Outside'}({list(self.cell)})"


def _f32(x):
    return np.float32(x)


class Grid:
    """grid.rs:30-170, with the reference's FP32 arithmetic (numpy float32
    scalars round every operation to single precision)."""

    def __init__(self, first_cell, cell_size, cell_count):
        self.first_cell = np.asarray(first_cell, dtype=np.float32).reshape(3).copy()
        self.cell_size = np.asarray(cell_size, dtype=np.float32).reshape(3).copy()
        self.cell_count = tuple(int(c) for c in cell_count)

    new = classmethod(lambda cls, first_cell, cell_size, cell_count: cls(first_cell, cell_size, cell_count))

    @classmethod
    def from_bounding_box(cls, bbox_min, bbox_max, cell_count):
        lo = np.asarray(bbox_min, dtype=np.float32).reshape(3)
        hi = np.asarray(bbox_max, dtype=np.float32).reshape(3)
        fc = np.array([_f32(c) for c in cell_count], dtype=np.float32)
        size = (hi - lo) / fc  # comp_div
        first = lo + size * np.float32(0.5)
        return cls(first, size, cell_count)

    def get_first_cell(self):
        return self.first_cell.copy()

    def get_last_cell(self):
        return np.array([self.first_cell[k] + _f32(self.cell_count[k]) * self.cell_size[k] for k in range(3)],
                        dtype=np.float32)

    def get_cell_size(self):
        return self.cell_size.copy()

    def get_cell_count(self):
        return list(self.cell_count)

    def get_total_cell_count(self):
        return self.cell_count[0] * self.cell_count[1] * self.cell_count[2]

    def get_bounding_box(self):
        mn = self.first_cell - self.cell_size * np.float32(0.5)
        mx = np.array([mn[k] + _f32(self.cell_count[k]) * self.cell_size[k] for k in range(3)], dtype=np.float32)
        return mn, mx

    def get_cell_idx(self, cell):
        return int(cell[2]) + int(cell[1]) * self.cell_count[2] + int(cell[0]) * self.cell_count[1] * self.cell_count[2]

    def get_cell_integer_coordinates(self, cell_idx):
        z = cell_idx % self.cell_count[2]
        y = (cell_idx // self.cell_count[2]) % self.cell_count[1]
        x = cell_idx // (self.cell_count[1] * self.cell_count[2])
        return [x, y, z]

    def get_cell_center(self, cell):
        return np.array([self.first_cell[k] + _f32(int(cell[k])) * self.cell_size[k] for k in range(3)],
                        dtype=np.float32)

    def snap_point_to_grid(self, point):
        p = np.asarray(point, dtype=np.float32).reshape(3)
        c = (p - self.get_bounding_box()[0]) / self.cell_size
        cell = []
        for k in range(3):
            f = np.floor(c[k])
            cell.append(0 if np.isnan(f) else int(max(min(f, 2.0 ** 63 - 1), -(2.0 ** 63))))
        ires = []
        for k in range(3):
            n = self.cell_count[k]
            if n == 0:
                raise SdfxError(_lib.ERR_INVALID_GRID, "assertion failed: min <= max")
            ires.append(min(max(cell[k], 0), n - 1))
        return SnapResult(ires == cell, ires)

    def cells_centers(self) -> np.ndarray:
        """All cell centres in output order (z fastest)."""
        n = self.cell_count
        xs = self.first_cell[0] + np.arange(n[0]).astype(np.float32) * self.cell_size[0]
        ys = self.first_cell[1] + np.arange(n[1]).astype(np.float32) * self.cell_size[1]
        zs = self.first_cell[2] + np.arange(n[2]).astype(np.float32) * self.cell_size[2]
        X, Y, Z = np.meshgrid(xs.astype(np.float32), ys.astype(np.float32), zs.astype(np.float32), indexing="ij")
        return np.stack([X, Y, Z], -1).reshape(-1, 3).astype(np.float32)

    def __eq__(self, other):
        return (isinstance(other, Grid) and np.array_equal(self.first_cell, other.first_cell)
                and np.array_equal(self.cell_size, other.cell_size) and self.cell_count == other.cell_count)

    def __repr__(self):
        return f"Grid(first_cell={self.first_cell.tolist()}, cell_size={self.cell_size.tolist()}, " \
               f"cell_count={list(self.cell_count)})"


def _verts(vertices) -> np.ndarray:
    return np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


def generate_sdf(vertices, indices: Topology, query_points, acceleration_method: _Accel = AccelerationMethod.RtreeBvh
                 ) -> np.ndarray:
    """lib.rs:364-391. Returns float32 signed distances (empty for RtreeBvh on an empty mesh)."""
    v = _verts(vertices)
    q = _verts(query_points)
    out = np.empty(len(q), dtype=np.float32)
    n = ctypes.c_uint64(0)
    idx = indices.indices
    _lib.check(_lib.lib().sdfx_generate_sdf(
        _ptr(v, _F), len(v), _ptr(idx, _U32), 0 if idx is None else idx.size, indices.kind, _ptr(q, _F), len(q),
        acceleration_method.kind, acceleration_method.sign, _ptr(out, _F), ctypes.byref(n)))
    return out[: n.value]


def _grid_call(fn, vertices, indices: Topology, grid: Grid, sign_method: int) -> np.ndarray:
    v = _verts(vertices)
    cnt = np.array(grid.cell_count, dtype=np.uint64)
    out = np.empty(grid.get_total_cell_count(), dtype=np.float32)
    idx = indices.indices
    first = np.ascontiguousarray(grid.first_cell, dtype=np.float32)
    size = np.ascontiguousarray(grid.cell_size, dtype=np.float32)
    _lib.check(fn(_ptr(v, _F), len(v), _ptr(idx, _U32), 0 if idx is None else idx.size, indices.kind,
                  _ptr(first, _F), _ptr(size, _F), _ptr(cnt, _U64), sign_method, _ptr(out, _F)))
    return out


def generate_grid_sdf(vertices, indices: Topology, grid: Grid, sign_method: int = SignMethod.Raycast) -> np.ndarray:
    """generate/grid.rs:265-273. Returns float32 distances, z fastest."""
    return _grid_call(_lib.lib().sdfx_generate_grid_sdf, vertices, indices, grid, sign_method)


def generate_grid_sdf_exact(vertices, indices: Topology, grid: Grid,
                            sign_method: int = SignMethod.Raycast) -> np.ndarray:
    """generate_grid_sdf without the flood fill's approximation (no counterpart
    in the reference): each cell holds the FP32 minimum over all triangles of
    the distance from its centre (Raycast, signed by the same grid raycasts as
    generate/grid.rs:568-684) or the compare_distances fold of the signed
    distances (Normal, as Bvh(Normal) answers a point, generic/bvh.rs:117-129).
    Same arguments and errors as generate_grid_sdf."""
    return _grid_call(_lib.lib().sdfx_generate_grid_sdf_exact, vertices, indices, grid, sign_method)


def generate_grid_sdf_batch(requests, sign_method: int = SignMethod.Raycast, return_status: bool = False):
    """Many generate_grid_sdf calls (generate/grid.rs:265-273) in one device run:
    `requests` is a sequence of (vertices, topology, grid). Each result is
    bit-identical to generate_grid_sdf on that request alone. Raises the first
    failing request's error, or with return_status=True returns
    (results, statuses) with None for failed requests."""
    reqs = (_lib.GridRequest * max(1, len(requests)))()
    keep = []
    outs = []
    for r, (vertices, indices, grid) in zip(reqs, requests):
        v = _verts(vertices)
        idx = None if indices.indices is None else np.ascontiguousarray(indices.indices, dtype=np.uint32).reshape(-1)
        out = np.empty(grid.get_total_cell_count(), dtype=np.float32)
        keep += [v, idx]
        outs.append(out)
        r.vertices = v.ctypes.data
        r.vertex_count = len(v)
        r.indices = None if idx is None else idx.ctypes.data
        r.index_count = 0 if idx is None else idx.size
        r.topology = indices.kind
        r.first_cell[:] = [float(x) for x in grid.first_cell]
        r.cell_size[:] = [float(x) for x in grid.cell_size]
        r.cell_count[:] = list(grid.cell_count)
        r.out = out.ctypes.data
    st = _lib.lib().sdfx_generate_grid_sdf_batch(reqs, len(requests), int(sign_method))
    if return_status:
        statuses = [reqs[i].status for i in range(len(requests))]
        return [o if s == _lib.OK else None for o, s in zip(outs, statuses)], statuses
    _lib.check(st)
    return outs


class SdfAccelerationMesh:
    """lib.rs:303-312: a device-resident mesh and its LBVH (freed on close/GC)."""

    def __init__(self, handle: int, method: _Accel):
        self._h = ctypes.c_void_p(handle)
        self.method = method

    def close(self):
        if self._h:
            _lib.lib().sdfx_free_sdf_acceleration(self._h)
            self._h = ctypes.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def build_sdf_acceleration(vertices, indices: Topology, acceleration_method: _Accel = AccelerationMethod.RtreeBvh
                           ) -> Optional[SdfAccelerationMesh]:
    """lib.rs:410-433. Returns None where the reference returns None."""
    v = _verts(vertices)
    idx = indices.indices
    h = ctypes.c_void_p(None)
    st = _lib.check(_lib.lib().sdfx_build_sdf_acceleration(
        _ptr(v, _F), len(v), _ptr(idx, _U32), 0 if idx is None else idx.size, indices.kind,
        acceleration_method.kind, acceleration_method.sign, ctypes.byref(h)))
    if st == _lib.EMPTY_RESULT or not h.value:
        return None
    return SdfAccelerationMesh(h.value, acceleration_method)


def query_sdf(acceleration: SdfAccelerationMesh, query_points) -> np.ndarray:
    """lib.rs:463-485."""
    q = _verts(query_points)
    out = np.empty(len(q), dtype=np.float32)
    n = ctypes.c_uint64(0)
    _lib.check(_lib.lib().sdfx_query_sdf(acceleration._h, _ptr(q, _F), len(q), _ptr(out, _F), ctypes.byref(n)))
    return out[: n.value]


def query_sdf_batch(requests, return_status: bool = False):
    """query_sdf (lib.rs:463-485) for many (SdfAccelerationMesh, query_points)
    pairs in one call (sdfx_query_sdf_batch_device): requests of one query
    mode share the same launches and each result is bit-identical to query_sdf
    on that pair alone. Points go to the device and results come back as
    numpy arrays. Raises the first failing request's error, or with
    return_status=True returns (results, statuses) with None for failed
    requests."""
    import torch
    dev = torch.device("cuda", torch.cuda.current_device())
    reqs = (_lib.PointRequest * max(1, len(requests)))()
    keep = []
    for r, (acc, pts) in zip(reqs, requests):
        q = torch.from_numpy(np.ascontiguousarray(pts, dtype=np.float32).reshape(-1, 3)).to(dev)
        out = torch.empty(q.shape[0], dtype=torch.float32, device=dev)
        keep.append((q, out))
        r.mesh = acc._h.value
        r.query_points = q.data_ptr() if q.numel() else None
        r.query_count = q.shape[0]
        r.out = out.data_ptr() if out.numel() else None
    stream = torch.cuda.current_stream()
    stream.synchronize()
    st = _lib.lib().sdfx_query_sdf_batch_device(reqs, len(requests), ctypes.c_void_p(stream.cuda_stream))
    statuses = [reqs[i].status for i in range(len(requests))]
    outs = [out.cpu().numpy() if s_ == _lib.OK else None for (q, out), s_ in zip(keep, statuses)]
    if return_status:
        return outs, statuses
    _lib.check(st)
    return outs


def query_vector_to_closest_point_rtree_bvh(acceleration: SdfAccelerationMesh, query_points):
    """generate/generic/rtree_bvh.rs:217-276: (closest - query) vectors and inside flags."""
    q = _verts(query_points)
    vec = np.empty((len(q), 3), dtype=np.float32)
    inside = np.empty(len(q), dtype=np.uint8)
    _lib.check(_lib.lib().sdfx_query_vector_to_closest_point(acceleration._h, _ptr(q, _F), len(q), _ptr(vec, _F),
                                                             _ptr(inside, _U8)))
    return vec, inside.astype(bool)


def check_mesh_triangle_list(vertices, faces) -> Optional[str]:
    """generate/mesh_check.rs:24-81 (re-exported at lib.rs:149), run on the
    device: None for the reference's Ok(()), the reference's Err message
    otherwise (duplicate vertex / face, duplicate or unmatched edge)."""
    v = _verts(vertices)
    f = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1)
    kind = ctypes.c_int(0)
    where = ctypes.c_uint64(0)
    msg = ctypes.create_string_buffer(512)
    st = _lib.check(_lib.lib().sdfx_check_mesh_triangle_list(_ptr(v, _F), len(v), _ptr(f, _U32), f.size,
                                                             ctypes.byref(kind), ctypes.byref(where), msg, 512))
    return None if st == _lib.OK else msg.value.decode()
