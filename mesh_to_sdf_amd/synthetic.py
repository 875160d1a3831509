"""Seeded synthetic meshes and query sets for the SDF workloads.

All generators are deterministic given their seed (numpy PCG64 streams) and
produce watertight, outward-wound (counter-clockwise seen from outside)
triangle lists with u32 indices:

* ``icosphere(level)``: icosahedron refined ``level`` times by 1->4 midpoint
  subdivision with an edge-midpoint cache, projected to the unit sphere
  (level 3: 642 vertices / 1,280 triangles; level 7: 163,842 / 327,680).
* ``torus(R, r, n_major, n_minor, jitter)``: each vertex displaced along its
  tube normal by N(0, jitter).
* ``noise_sphere(level, amplitude)``: icosphere with radius 1 + a*n(p), n a
  3-octave value noise normalised to [-1, 1].

Query sets: uniform points in the mesh AABB with half-extents scaled by 1.1
about its centre, and near-surface points (area-weighted triangle, uniform
barycentric point, offset along the face normal by N(0, sigma)).
"""
from __future__ import annotations

import math

import numpy as np

SEED_BASE = 0x5DF00000


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def icosphere(level: int):
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [
        (-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0),
        (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
        (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1),
    ]
    faces = [
        (0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
        (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
        (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
        (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1),
    ]
    v = np.array(verts, dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array(faces, dtype=np.int64)
    for _ in range(level):
        # unique edges -> midpoint vertex ids
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=0)
        e_sorted = np.sort(e, axis=1)
        uniq, inv = np.unique(e_sorted, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mids = v[uniq[:, 0]] + v[uniq[:, 1]]
        mids /= np.linalg.norm(mids, axis=1, keepdims=True)
        base = len(v)
        v = np.concatenate([v, mids], axis=0)
        nf = len(f)
        m01 = base + inv[0:nf]
        m12 = base + inv[nf:2 * nf]
        m20 = base + inv[2 * nf:3 * nf]
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        f = np.stack([
            np.stack([a, m01, m20], 1),
            np.stack([b, m12, m01], 1),
            np.stack([c, m20, m12], 1),
            np.stack([m01, m12, m20], 1),
        ], axis=1).reshape(-1, 3)
    return v.astype(np.float32), f.astype(np.uint32)


def torus(R=1.0, r=0.35, n_major=256, n_minor=128, jitter=0.002, seed=SEED_BASE + 2):
    i = np.arange(n_major)
    j = np.arange(n_minor)
    phi = 2.0 * np.pi * i / n_major
    theta = 2.0 * np.pi * j / n_minor
    P, T = np.meshgrid(phi, theta, indexing="ij")
    nx = np.cos(T) * np.cos(P)
    ny = np.cos(T) * np.sin(P)
    nz = np.sin(T)
    x = (R + r * np.cos(T)) * np.cos(P)
    y = (R + r * np.cos(T)) * np.sin(P)
    z = r * np.sin(T)
    if jitter:
        eps = _rng(seed).normal(0.0, jitter, size=P.shape)
        x, y, z = x + nx * eps, y + ny * eps, z + nz * eps
    v = np.stack([x, y, z], -1).reshape(-1, 3)
    I, J = np.meshgrid(i, j, indexing="ij")
    v00 = I * n_minor + J
    v10 = ((I + 1) % n_major) * n_minor + J
    v11 = ((I + 1) % n_major) * n_minor + (J + 1) % n_minor
    v01 = I * n_minor + (J + 1) % n_minor
    f = np.stack([
        np.stack([v00, v10, v11], -1),
        np.stack([v00, v11, v01], -1),
    ], axis=2).reshape(-1, 3)
    return v.astype(np.float32), f.astype(np.uint32)


def _hash3(ix, iy, iz, seed):
    h = (ix.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
         ^ iy.astype(np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F)
         ^ iz.astype(np.uint64) * np.uint64(0x165667B19E3779F9)
         ^ np.uint64(seed))
    h ^= h >> np.uint64(33)
    h *= np.uint64(0xFF51AFD7ED558CCD)
    h ^= h >> np.uint64(33)
    h *= np.uint64(0xC4CEB9FE1A85EC53)
    h ^= h >> np.uint64(33)
    return (h >> np.uint64(11)).astype(np.float64) / float(1 << 53) * 2.0 - 1.0


def value_noise(p: np.ndarray, seed: int, octaves=3, f0=2.0) -> np.ndarray:
    """Sum of `octaves` trilinear value-noise octaves (frequency f0*2^o,
    amplitude 2^-o), normalised to [-1, 1]."""
    total = np.zeros(len(p))
    norm = 0.0
    with np.errstate(over="ignore"):
        for o in range(octaves):
            f = f0 * (2 ** o)
            q = p * f + 1000.0
            i0 = np.floor(q).astype(np.int64)
            t = q - i0
            s = t * t * (3 - 2 * t)
            acc = np.zeros(len(p))
            for dx in (0, 1):
                for dy in (0, 1):
                    for dz in (0, 1):
                        w = ((s[:, 0] if dx else 1 - s[:, 0]) * (s[:, 1] if dy else 1 - s[:, 1])
                             * (s[:, 2] if dz else 1 - s[:, 2]))
                        acc += w * _hash3(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz, seed + o)
            total += acc * (0.5 ** o)
            norm += 0.5 ** o
    return total / norm


def noise_sphere(level=7, amplitude=0.05, seed=SEED_BASE + 3):
    v, f = icosphere(level)
    vd = v.astype(np.float64)
    n = value_noise(vd, seed)
    vd = vd * (1.0 + amplitude * n)[:, None]
    return vd.astype(np.float32), f


_ICOSPHERES: dict = {}


def jittered_icosphere(level: int, jitter=0.002, seed=SEED_BASE):
    """Icosphere with each vertex's radius scaled by 1 + N(0, jitter)."""
    if level not in _ICOSPHERES:
        _ICOSPHERES[level] = icosphere(level)
    v, f = _ICOSPHERES[level]
    eps = _rng(seed).normal(0.0, jitter, size=len(v))
    return (v.astype(np.float64) * (1.0 + eps)[:, None]).astype(np.float32), f.copy()


def aabb(v: np.ndarray):
    return v.min(axis=0), v.max(axis=0)


def expanded_aabb(v: np.ndarray, factor=1.1):
    lo, hi = aabb(v.astype(np.float64))
    c = 0.5 * (lo + hi)
    h = 0.5 * (hi - lo) * factor
    return (c - h).astype(np.float32), (c + h).astype(np.float32)


def bbox_diagonal(v: np.ndarray) -> float:
    lo, hi = aabb(v.astype(np.float64))
    return float(np.linalg.norm(hi - lo))


def uniform_queries(v: np.ndarray, n: int, seed: int, factor=1.1) -> np.ndarray:
    lo, hi = expanded_aabb(v, factor)
    u = _rng(seed).random((n, 3))
    return (lo + u * (hi - lo)).astype(np.float32)


def surface_queries(v: np.ndarray, f: np.ndarray, n: int, sigma: float, seed: int) -> np.ndarray:
    g = _rng(seed)
    a, b, c = (v[f[:, k]].astype(np.float64) for k in range(3))
    nrm = np.cross(b - a, c - a)
    area = np.linalg.norm(nrm, axis=1)
    tri = g.choice(len(f), size=n, p=area / area.sum())
    r1 = np.sqrt(g.random(n))
    r2 = g.random(n)
    w0, w1, w2 = 1 - r1, r1 * (1 - r2), r1 * r2
    p = a[tri] * w0[:, None] + b[tri] * w1[:, None] + c[tri] * w2[:, None]
    nn = nrm[tri] / np.maximum(area[tri], 1e-30)[:, None]
    p += nn * g.normal(0.0, sigma, size=n)[:, None]
    return p.astype(np.float32)


def grid_for(v: np.ndarray, counts=(128, 128, 128), factor=1.1):
    """Grid::from_bounding_box over the expanded AABB (grid.rs:213-228)."""
    lo, hi = expanded_aabb(v, factor)
    cnt = np.array(counts, dtype=np.float32)
    size = ((hi - lo) / cnt).astype(np.float32)
    first = (lo + size * np.float32(0.5)).astype(np.float32)
    return first, size, tuple(int(c) for c in counts)


# ---- the five workloads ----

def config1(seed=SEED_BASE + 1):
    """Unit icosphere L3 (642 V / 1,280 T), 10,000 uniform queries; RtreeBvh."""
    v, f = icosphere(3)
    q = uniform_queries(v, 10_000, seed)
    return v, f, q


def config2():
    """Jittered torus 256x128 (65,536 T), 128^3 grid from the expanded AABB; Raycast."""
    v, f = torus()
    first, size, count = grid_for(v, (128, 128, 128))
    return v, f, first, size, count


def config3(n_queries=4_194_304, level=7, seed=SEED_BASE + 3):
    """Noise-displaced sphere (327,680 T at level 7); half near-surface, half uniform queries; Bvh(Raycast)."""
    v, f = noise_sphere(level, 0.05, seed)
    diag = bbox_diagonal(v)
    n_s = n_queries // 2
    qs = surface_queries(v, f, n_s, 0.01 * diag, seed + 100)
    qu = uniform_queries(v, n_queries - n_s, seed + 200)
    return v, f, np.concatenate([qs, qu], axis=0)


def config4(n_meshes=256, seed=SEED_BASE + 4, t_min=1_280, t_max=327_680, edge_min=32, edge_max=128):
    """The batched workload: `n_meshes` watertight meshes, alternately jittered
    icospheres and jittered tori, with triangle counts log-uniform in
    [t_min, t_max] (the icosphere level, 20 * 4^level triangles, or the torus
    segment counts 2N x N, 4N^2 triangles, nearest to the drawn target), each
    with one grid whose axes are independently log-uniform in
    [edge_min, edge_max] cells over its AABB expanded by 10%. Returns a list of
    (vertices, faces, first_cell, cell_size, cell_count)."""
    g = _rng(seed)
    out = []
    for i in range(n_meshes):
        target = math.exp(g.uniform(math.log(t_min), math.log(t_max)))
        mseed = seed + 1000 + i
        if i % 2 == 0:
            level = min(range(3, 8), key=lambda lv: abs(math.log(20 * 4 ** lv) - math.log(target)))
            v, f = jittered_icosphere(level, 0.002, mseed)
        else:
            n = max(3, int(round(math.sqrt(target / 4.0))))
            v, f = torus(1.0, float(g.uniform(0.25, 0.45)), 2 * n, n, 0.002, mseed)
        counts = tuple(int(round(math.exp(g.uniform(math.log(edge_min), math.log(edge_max))))) for _ in range(3))
        first, size, count = grid_for(v, counts)
        out.append((v, f, first, size, count))
    return out


def serve_requests(meshes, n_requests=256, seed=SEED_BASE + 6, point_min=1_000, point_max=1_000_000,
                   edge_min=32, edge_max=128):
    """The served workload's request stream over `meshes` (a list of
    (vertices, faces, ...) as config4 returns): requests alternate point
    queries and grids, each on a mesh drawn uniformly. A point request holds n
    points, n log-uniform in [point_min, point_max], half near the surface
    (normal offsets N(0, 0.01 * bbox diagonal)) and half uniform in the AABB
    expanded by 10%; a grid request has e^3 cells, e log-uniform in
    [edge_min, edge_max], over the AABB expanded by 10%. Returns a list of
    ("points", mesh, points) and ("grid", mesh, (first_cell, cell_size,
    cell_count)); the points are generated here, so large streams take memory
    (about 12 bytes per point)."""
    g = _rng(seed)
    out = []
    for k in range(n_requests):
        mi = int(g.integers(len(meshes)))
        v, f = meshes[mi][0], meshes[mi][1]
        if k % 2 == 0:
            n = int(round(math.exp(g.uniform(math.log(point_min), math.log(point_max)))))
            q = np.concatenate([surface_queries(v, f, n // 2, 0.01 * bbox_diagonal(v), seed + 10 + k),
                                uniform_queries(v, n - n // 2, seed + 20 + k)])
            out.append(("points", mi, q))
        else:
            e = int(round(math.exp(g.uniform(math.log(edge_min), math.log(edge_max)))))
            out.append(("grid", mi, grid_for(v, (e, e, e))))
    return out
