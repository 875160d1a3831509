"""The served form: meshes resident on the device and a queue of point and grid
requests answered by libsdfx's server (include/sdfx.h, sdfx_server_*).

    srv = SdfServer(hbm_budget_bytes=8 << 30, workers=2)
    mid = srv.add_mesh(vertices, Topology.TriangleList(faces))
    t1 = srv.submit_points(mid, points, AccelerationMethod.Bvh(SignMethod.Raycast))
    t2 = srv.submit_grid(mid, Grid(first, size, count), SignMethod.Raycast)
    d1, ms1 = srv.wait(t1)      # device tensors with the answers, latency in ms
    d2, ms2 = srv.wait(t2)

Requests submitted while the device is busy are batched with the other queued
requests when a worker comes free (continuous batching); each answer is
bit-identical to the request served alone. Query points and outputs live in
device memory (torch tensors on the server's device); host arrays are copied
there at submission.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib, Grid, SignMethod, Topology
from ._lib import SdfxError

__all__ = ["SdfServer", "Ticket"]


class Ticket:
    """A submitted request: its id and the device buffers it uses."""

    def __init__(self, ticket: int, out, keep):
        self.ticket = ticket
        self.out = out
        self._keep = keep  # inputs that must outlive the request


class SdfServer:
    def __init__(self, hbm_budget_bytes: int, workers: int = 2):
        import torch
        self._torch = torch
        self.device = torch.device("cuda", torch.cuda.current_device())
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().sdfx_server_create(int(hbm_budget_bytes), int(workers), ctypes.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            _lib.lib().sdfx_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _dev(self, a, dtype):
        t = self._torch
        if isinstance(a, t.Tensor):
            return a.to(device=self.device, dtype=dtype).contiguous()
        return t.from_numpy(np.ascontiguousarray(a)).to(device=self.device, dtype=dtype).contiguous()

    def add_mesh(self, vertices, indices: Topology) -> int:
        """Uploads the mesh and builds its LBVH; returns its id (SdfxError with
        ERR_OVER_BUDGET when it cannot fit the budget)."""
        v = self._dev(np.asarray(vertices, dtype=np.float32).reshape(-1, 3), self._torch.float32)
        idx = None
        if indices.indices is not None:
            idx = self._dev(np.asarray(indices.indices, dtype=np.uint32).view(np.int32), self._torch.int32)
        mid = ctypes.c_uint32(0)
        _lib.check(_lib.lib().sdfx_server_add_mesh(
            self._h, v.data_ptr(), v.shape[0], None if idx is None else idx.data_ptr(),
            0 if idx is None else idx.numel(), indices.kind, ctypes.byref(mid)))
        self._torch.cuda.synchronize()
        return int(mid.value)

    def add_meshes(self, meshes) -> list:
        """Adds (vertices, topology) pairs with one batched LBVH build; returns
        their ids (raises the first failing mesh's error)."""
        t = self._torch
        descs = (_lib.ServerMesh * max(1, len(meshes)))()
        keep = []
        for d, (vertices, indices) in zip(descs, meshes):
            v = self._dev(np.asarray(vertices, dtype=np.float32).reshape(-1, 3), t.float32)
            idx = None
            if indices.indices is not None:
                idx = self._dev(np.asarray(indices.indices, dtype=np.uint32).view(np.int32), t.int32)
            keep += [v, idx]
            d.vertices, d.vertex_count = v.data_ptr() if v.numel() else None, v.shape[0]
            d.indices = None if idx is None or not idx.numel() else idx.data_ptr()
            d.index_count = 0 if idx is None else idx.numel()
            d.topology = indices.kind
        _lib.check(_lib.lib().sdfx_server_add_meshes(self._h, descs, len(meshes)))
        t.cuda.synchronize()
        return [int(descs[i].id) for i in range(len(meshes))]

    def remove_mesh(self, mesh_id: int) -> None:
        _lib.check(_lib.lib().sdfx_server_remove_mesh(self._h, int(mesh_id)))

    def submit_points(self, mesh_id: int, points, method, out=None) -> Ticket:
        t = self._torch
        q = self._dev(points, t.float32).reshape(-1, 3)
        if out is None:
            out = t.empty(q.shape[0], dtype=t.float32, device=self.device)
        t.cuda.current_stream().synchronize()  # the server's streams read q
        tk = ctypes.c_uint64(0)
        _lib.check(_lib.lib().sdfx_server_submit_points(self._h, int(mesh_id), q.data_ptr(), q.shape[0],
                                                        method.kind, method.sign, out.data_ptr(), ctypes.byref(tk)))
        return Ticket(int(tk.value), out, q)

    def submit_grid(self, mesh_id: int, grid: Grid, sign_method: int = SignMethod.Raycast, out=None) -> Ticket:
        t = self._torch
        if out is None:
            out = t.empty(grid.get_total_cell_count(), dtype=t.float32, device=self.device)
        t.cuda.current_stream().synchronize()
        first = (ctypes.c_float * 3)(*[float(x) for x in grid.first_cell])
        size = (ctypes.c_float * 3)(*[float(x) for x in grid.cell_size])
        count = (ctypes.c_uint64 * 3)(*[int(x) for x in grid.cell_count])
        tk = ctypes.c_uint64(0)
        _lib.check(_lib.lib().sdfx_server_submit_grid(self._h, int(mesh_id), first, size, count, int(sign_method),
                                                      out.data_ptr() if out.numel() else None, ctypes.byref(tk)))
        return Ticket(int(tk.value), out, None)

    def wait(self, ticket: Ticket):
        """(answers as a device tensor, latency in ms); None for the answers when
        the reference returns an empty Vec. Raises the request's error."""
        ms = ctypes.c_double(0.0)
        st = _lib.lib().sdfx_server_wait(self._h, ticket.ticket, ctypes.byref(ms))
        if st < 0:
            raise SdfxError(st, _lib.lib().sdfx_last_error().decode())
        return (None if st == _lib.EMPTY_RESULT else ticket.out), ms.value

    def cancel(self, ticket: Ticket) -> bool:
        """Cancels the request if it is still queued (its wait then raises
        ERR_CANCELLED); False when it already runs or is done."""
        c = ctypes.c_int(0)
        _lib.check(_lib.lib().sdfx_server_cancel(self._h, ticket.ticket, ctypes.byref(c)))
        return bool(c.value)

    def done(self, ticket: Ticket) -> bool:
        d = ctypes.c_int(0)
        _lib.check(_lib.lib().sdfx_server_poll(self._h, ticket.ticket, ctypes.byref(d)))
        return bool(d.value)

    def stats(self) -> dict:
        s = _lib.ServerStats()
        _lib.check(_lib.lib().sdfx_server_get_stats(self._h, ctypes.byref(s)))
        return {name: int(getattr(s, name)) for name, _ in _lib.ServerStats._fields_}
