"""The reference's on-disk SDF format (serde.rs:77-221), so results computed here
load in existing consumers and files they wrote load here.

The reference serialises `SerializeVersion::V1(&SerializeSdf)` with
`rmp_serde::to_vec` (serde.rs:160-165): MessagePack with enums as one-key maps
and structs as arrays, f32 as MessagePack float32 and usize as the shortest
unsigned integer. The two payloads are

    {"V1": {"Generic": [query_points, distances]}}          (SerializeGeneric)
    {"V1": {"Grid": [[first_cell, cell_size, cell_count], distances]}}  (SerializeGrid)

with points as 3-element float32 arrays. This module writes exactly those bytes
(numpy builds the float runs, so an 8 MiB grid writes in milliseconds) and
reads them back, also accepting float64 and any integer width where the
reference's decoder would.
"""
from __future__ import annotations

import os
import struct
from dataclasses import dataclass

import numpy as np

from . import Grid

__all__ = ["SerdeError", "SerializeGeneric", "SerializeGrid", "SerializeSdf", "DeserializeGeneric",
           "DeserializeGrid", "DeserializeSdf", "serialize", "deserialize", "save_to_file", "read_from_file"]


class SerdeError(Exception):
    """serde.rs:44-52: SerializationFailed, DeserializationFailed or IoError."""

    SERIALIZATION_FAILED = "SerializationFailed"
    DESERIALIZATION_FAILED = "DeserializationFailed"
    IO_ERROR = "IoError"

    def __init__(self, kind: str, detail: str = ""):
        super().__init__(f"{kind}: {detail}" if detail else kind)
        self.kind = kind


@dataclass
class SerializeGeneric:
    """serde.rs:88-94: query points and their distances (generate_sdf)."""
    query_points: object
    distances: object


@dataclass
class SerializeGrid:
    """serde.rs:99-105: a grid and its cell distances (generate_grid_sdf)."""
    grid: Grid
    distances: object


class SerializeSdf:
    """serde.rs:77-83: `SerializeSdf::Generic(..)` / `SerializeSdf::Grid(..)`."""
    Generic = SerializeGeneric
    Grid = SerializeGrid


@dataclass
class DeserializeGeneric:
    """serde.rs:130-136."""
    query_points: np.ndarray  # (n, 3) float32
    distances: np.ndarray     # (n,) float32


@dataclass
class DeserializeGrid:
    """serde.rs:141-147."""
    grid: Grid
    distances: np.ndarray


class DeserializeSdf:
    """serde.rs:120-125."""
    Generic = DeserializeGeneric
    Grid = DeserializeGrid


# ---------------------------------------------------------------- writing

def _fixstr(s: str) -> bytes:
    b = s.encode()
    assert len(b) < 32
    return bytes([0xA0 | len(b)]) + b


def _array_header(n: int) -> bytes:
    if n < 16:
        return bytes([0x90 | n])
    if n < 1 << 16:
        return b"\xdc" + struct.pack(">H", n)
    if n < 1 << 32:
        return b"\xdd" + struct.pack(">I", n)
    raise SerdeError(SerdeError.SERIALIZATION_FAILED, f"array of {n} elements")


def _uint(v: int) -> bytes:
    if v < 0:
        raise SerdeError(SerdeError.SERIALIZATION_FAILED, f"negative cell count {v}")
    if v < 128:
        return bytes([v])
    if v < 1 << 8:
        return b"\xcc" + struct.pack(">B", v)
    if v < 1 << 16:
        return b"\xcd" + struct.pack(">H", v)
    if v < 1 << 32:
        return b"\xce" + struct.pack(">I", v)
    return b"\xcf" + struct.pack(">Q", v)


def _f32_run(values: np.ndarray) -> bytes:
    """float32 MessagePack elements (0xca + big-endian bits) for each value."""
    v = np.ascontiguousarray(values, dtype=np.float32).reshape(-1)
    rec = np.empty(v.size, dtype=[("tag", "u1"), ("bits", ">f4")])
    rec["tag"] = 0xCA
    rec["bits"] = v
    return rec.tobytes()


def _points(points) -> bytes:
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    rec = np.empty(p.shape[0], dtype=[("hdr", "u1"), ("t0", "u1"), ("x", ">f4"), ("t1", "u1"), ("y", ">f4"),
                                      ("t2", "u1"), ("z", ">f4")])
    rec["hdr"] = 0x93
    rec["t0"] = rec["t1"] = rec["t2"] = 0xCA
    rec["x"], rec["y"], rec["z"] = p[:, 0], p[:, 1], p[:, 2]
    return _array_header(p.shape[0]) + rec.tobytes()


def serialize(sdf) -> bytes:
    """serde.rs:160-165 (rmp_serde::to_vec of SerializeVersion::V1(sdf))."""
    head = b"\x81" + _fixstr("V1") + b"\x81"
    if isinstance(sdf, SerializeGeneric):
        d = np.ascontiguousarray(sdf.distances, dtype=np.float32).reshape(-1)
        return head + _fixstr("Generic") + b"\x92" + _points(sdf.query_points) + _array_header(d.size) + _f32_run(d)
    if isinstance(sdf, SerializeGrid):
        g = sdf.grid
        d = np.ascontiguousarray(sdf.distances, dtype=np.float32).reshape(-1)
        grid = (b"\x93" + b"\x93" + _f32_run(g.get_first_cell()) + b"\x93" + _f32_run(g.get_cell_size()) + b"\x93"
                + b"".join(_uint(int(c)) for c in g.get_cell_count()))
        return head + _fixstr("Grid") + b"\x92" + grid + _array_header(d.size) + _f32_run(d)
    raise SerdeError(SerdeError.SERIALIZATION_FAILED, f"not a SerializeSdf value: {type(sdf).__name__}")


def save_to_file(sdf, path) -> None:
    """serde.rs:186-192."""
    data = serialize(sdf)
    try:
        with open(os.fspath(path), "wb") as fh:
            fh.write(data)
    except OSError as e:
        raise SerdeError(SerdeError.IO_ERROR, str(e)) from e


# ---------------------------------------------------------------- reading

class _Reader:
    def __init__(self, data: bytes):
        self.b = memoryview(data)
        self.i = 0

    def fail(self, what):
        raise SerdeError(SerdeError.DESERIALIZATION_FAILED, f"{what} at byte {self.i}")

    def byte(self) -> int:
        if self.i >= len(self.b):
            self.fail("unexpected end of data")
        v = self.b[self.i]
        self.i += 1
        return v

    def take(self, n: int) -> bytes:
        if self.i + n > len(self.b):
            self.fail("unexpected end of data")
        v = bytes(self.b[self.i:self.i + n])
        self.i += n
        return v

    def map_len(self) -> int:
        t = self.byte()
        if 0x80 <= t <= 0x8F:
            return t & 0x0F
        if t == 0xDE:
            return struct.unpack(">H", self.take(2))[0]
        if t == 0xDF:
            return struct.unpack(">I", self.take(4))[0]
        self.fail("expected a map")

    def array_len(self) -> int:
        t = self.byte()
        if 0x90 <= t <= 0x9F:
            return t & 0x0F
        if t == 0xDC:
            return struct.unpack(">H", self.take(2))[0]
        if t == 0xDD:
            return struct.unpack(">I", self.take(4))[0]
        self.fail("expected an array")

    def string(self) -> str:
        t = self.byte()
        if 0xA0 <= t <= 0xBF:
            n = t & 0x1F
        elif t == 0xD9:
            n = self.byte()
        elif t == 0xDA:
            n = struct.unpack(">H", self.take(2))[0]
        elif t == 0xDB:
            n = struct.unpack(">I", self.take(4))[0]
        else:
            self.fail("expected a string")
        return self.take(n).decode()

    def variant(self) -> str:
        """An enum variant: rmp_serde writes `{"Name": value}`."""
        if self.map_len() != 1:
            self.fail("expected a one-entry map (enum variant)")
        return self.string()

    def number(self) -> float:
        t = self.byte()
        if t == 0xCA:
            return struct.unpack(">f", self.take(4))[0]
        if t == 0xCB:
            return struct.unpack(">d", self.take(8))[0]
        if t <= 0x7F:
            return float(t)
        if t >= 0xE0:
            return float(t - 0x100)
        fmt = {0xCC: ">B", 0xCD: ">H", 0xCE: ">I", 0xCF: ">Q", 0xD0: ">b", 0xD1: ">h", 0xD2: ">i", 0xD3: ">q"}.get(t)
        if fmt is None:
            self.fail("expected a number")
        return float(struct.unpack(fmt, self.take(struct.calcsize(fmt)))[0])

    def uint(self) -> int:
        t = self.byte()
        if t <= 0x7F:
            return t
        fmt = {0xCC: ">B", 0xCD: ">H", 0xCE: ">I", 0xCF: ">Q"}.get(t)
        if fmt is None:
            self.fail("expected an unsigned integer")
        return struct.unpack(fmt, self.take(struct.calcsize(fmt)))[0]

    def f32_array(self, n: int) -> np.ndarray:
        # fast path: n float32 elements back to back
        end = self.i + 5 * n
        if end <= len(self.b):
            raw = np.frombuffer(self.b[self.i:end], dtype=[("tag", "u1"), ("bits", ">f4")])
            if n == 0 or np.all(raw["tag"] == 0xCA):
                self.i = end
                return raw["bits"].astype(np.float32)
        return np.array([self.number() for _ in range(n)], dtype=np.float32)

    def point(self) -> np.ndarray:
        if self.array_len() != 3:
            self.fail("expected a 3-element point")
        return self.f32_array(3)


def deserialize(data: bytes):
    """serde.rs:168-176: reads a V1 payload."""
    r = _Reader(bytes(data))
    if r.variant() != "V1":
        r.fail("unknown version")
    kind = r.variant()
    if r.array_len() != 2:
        r.fail("expected a 2-field struct")
    if kind == "Generic":
        n = r.array_len()
        pts = np.empty((n, 3), dtype=np.float32)
        end = r.i + 16 * n
        fast = False
        if end <= len(r.b) and n:
            raw = np.frombuffer(r.b[r.i:end], dtype=[("hdr", "u1"), ("t0", "u1"), ("x", ">f4"), ("t1", "u1"),
                                                      ("y", ">f4"), ("t2", "u1"), ("z", ">f4")])
            if np.all(raw["hdr"] == 0x93) and np.all(raw["t0"] == 0xCA) and np.all(raw["t1"] == 0xCA) \
                    and np.all(raw["t2"] == 0xCA):
                pts[:, 0], pts[:, 1], pts[:, 2] = raw["x"], raw["y"], raw["z"]
                r.i = end
                fast = True
        if not fast:
            for k in range(n):
                pts[k] = r.point()
        m = r.array_len()
        out = DeserializeGeneric(pts, r.f32_array(m))
    elif kind == "Grid":
        if r.array_len() != 3:
            r.fail("expected a 3-field Grid")
        first, size = r.point(), r.point()
        if r.array_len() != 3:
            r.fail("expected 3 cell counts")
        count = [r.uint() for _ in range(3)]
        m = r.array_len()
        out = DeserializeGrid(Grid(first, size, count), r.f32_array(m))
    else:
        r.fail(f"unknown variant {kind!r}")
    return out  # like rmp_serde::from_slice, bytes after the value are not read


def read_from_file(path):
    """serde.rs:214-220."""
    try:
        with open(os.fspath(path), "rb") as fh:
            data = fh.read()
    except OSError as e:
        raise SerdeError(SerdeError.IO_ERROR, str(e)) from e
    return deserialize(data)
