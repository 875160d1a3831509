"""The schedule-free grid (generate_grid_sdf_exact): each cell's distance is
the FP32 minimum over all triangles, which the reference's flood fill
(generate/grid.rs:495-558) only approaches.

The oracle's restatement (GRID_EXACT) is pinned on the CPU against paths that
share no code with it: brute-force nearest distances, the None path's fold for
Normal, FP64 distances, and the frontier schedule's raycast signs. The GPU
tests hold the device to the oracle bit for bit."""
import numpy as np
import pytest

from mesh_to_sdf_amd import synthetic as S

import refcases as R


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def _meshes(suzanne):
    sv, sf = suzanne
    return [("torus", *S.torus(1.0, 0.35, 48, 24, 0.002)),
            ("noise sphere", *S.noise_sphere(3, 0.05)),
            ("tetrahedron", np.asarray(R.TETRA_VERTICES, np.float32), np.asarray(R.TETRA_INDICES, np.uint32)),
            ("suzanne", np.asarray(sv, np.float32), np.asarray(sf, np.uint32))]


@pytest.mark.parametrize("sign", [0, 1])
def test_exact_oracle_is_pinned(oracle, suzanne, sign):
    """Raycast: |value| is the square root of the brute-force minimum d^2 (the
    reference's point_triangle_distance to the nearest triangle), the signs
    are the frontier schedule's (the same grid raycasts). Normal: the value
    equals the None path's compare_distances fold at the cell centre
    (default.rs:22-85), which folds every triangle in order. Both: within
    1e-6 x bbox diagonal of an FP64 distance, and never above the frontier or
    the reference 1-heap schedule's magnitude."""
    O = oracle
    for name, v, f in _meshes(suzanne):
        f = np.asarray(f, np.uint32).reshape(-1)
        first, size, count = S.grid_for(v, (21, 17, 24))
        s, ex, _ = O.generate_grid_sdf(v, f, O.TOPO_LIST, first, size, count, sign, 4, O.GRID_EXACT)
        assert s == 0
        cells = O.grid_centres(first, size, count)
        s, fr, _ = O.generate_grid_sdf(v, f, O.TOPO_LIST, first, size, count, sign, 4, O.GRID_FRONTIER)
        s, h1, _ = O.generate_grid_sdf(v, f, O.TOPO_LIST, first, size, count, sign, 1, O.GRID_SEQUENTIAL)
        if sign == O.SIGN_RAYCAST:
            d2, _ = O.nearest_bruteforce(v, f.reshape(-1, 3), cells)
            assert np.array_equal(bits(np.abs(ex)), bits(np.sqrt(d2))), name
            assert np.array_equal(np.signbit(ex), np.signbit(fr)), name
        else:
            s, none = O.generate_sdf(v, f, O.TOPO_LIST, cells, O.ACCEL_NONE, O.SIGN_NORMAL)
            assert s == 0
            assert np.array_equal(bits(ex), bits(none)), name
        d64 = O.fp64_distance(v, f.reshape(-1, 3), cells)
        assert np.max(np.abs(np.abs(ex).astype(np.float64) - d64)) <= 1e-6 * S.bbox_diagonal(v), name
        assert np.all(np.abs(ex) <= np.abs(fr)) and np.all(np.abs(ex) <= np.abs(h1)), name


def test_exact_oracle_config2(oracle):
    """The bench grid (65,536 triangles, 128^3): every cell within 1e-5 x bbox
    diagonal of the exact distance by construction (it is that distance), the
    signs the frontier schedule's."""
    O = oracle
    v, f, first, size, count = S.config2()
    s, ex, _ = O.generate_grid_sdf(v, f, O.TOPO_LIST, first, size, count, O.SIGN_RAYCAST, O.threads(),
                                   O.GRID_EXACT)
    assert s == 0
    d2, _ = O.nearest_bvh(v, f, O.grid_centres(first, size, count))
    assert np.array_equal(bits(np.abs(ex)), bits(np.sqrt(d2)))
    s, fr, _ = O.generate_grid_sdf(v, f, O.TOPO_LIST, first, size, count, O.SIGN_RAYCAST, O.threads(),
                                   O.GRID_FRONTIER)
    assert np.array_equal(np.signbit(ex), np.signbit(fr))


@pytest.mark.gpu
@pytest.mark.parametrize("sign", [0, 1])
def test_exact_grid_matches_oracle(sdf, oracle, suzanne, sign):
    """Device against the oracle's GRID_EXACT, bit for bit, on a torus, a
    noise sphere, the reference's tetrahedron and suzanne, with uneven and
    one-cell axes and a negative cell size; an empty mesh and a zero-axis grid
    behave as generate_grid_sdf does."""
    O = oracle
    for name, v, f in _meshes(suzanne):
        f = np.asarray(f, np.uint32).reshape(-1)
        topo = sdf.Topology.TriangleList(f)
        lo, hi = v.min(0), v.max(0)
        grids = [S.grid_for(v, (21, 17, 24)), S.grid_for(v, (1, 9, 33))]
        first, size, count = S.grid_for(v, (12, 14, 10))
        first = np.array(first, np.float32)
        size = np.array(size, np.float32)
        first[1] = hi[1] + size[1] * np.float32(0.5)
        size[1] = -size[1]
        grids.append((first, size, count))
        for first, size, count in grids:
            out = sdf.generate_grid_sdf_exact(v, topo, sdf.Grid(first, size, count), sign)
            s, ref, _ = O.generate_grid_sdf(v, f, O.TOPO_LIST, first, size, count, sign, O.threads(), O.GRID_EXACT)
            assert s == 0
            bad = np.flatnonzero(bits(out) != bits(ref))
            assert bad.size == 0, (name, count, sign, bad[:5], out[bad[:5]], ref[bad[:5]])
    empty = sdf.generate_grid_sdf_exact(np.zeros((0, 3), np.float32), sdf.Topology.TriangleList(
        np.zeros(0, np.uint32)), sdf.Grid([0, 0, 0], [1, 1, 1], [2, 3, 4]), sign)
    assert np.all(empty == np.float32(np.finfo(np.float32).max))
    v, f = S.torus(1.0, 0.35, 16, 8, 0.002)
    with pytest.raises(sdf.SdfxError) as e:
        sdf.generate_grid_sdf_exact(v, sdf.Topology.TriangleList(f.reshape(-1)), sdf.Grid([0, 0, 0], [1, 1, 1],
                                                                                          [2, 0, 4]), sign)
    with pytest.raises(sdf.SdfxError) as e2:
        sdf.generate_grid_sdf(v, sdf.Topology.TriangleList(f.reshape(-1)), sdf.Grid([0, 0, 0], [1, 1, 1], [2, 0, 4]),
                              sign)
    assert e.value.status == e2.value.status


@pytest.mark.gpu
def test_exact_grid_config2(sdf, oracle):
    """The bench grid at full size: every |value| is the square root of the
    nearest d^2 (the oracle's BVH restates the brute-force minimum), and the
    signs equal the flood-fill grid's (the same raycasts)."""
    O = oracle
    v, f, first, size, count = S.config2()
    topo = sdf.Topology.TriangleList(f.reshape(-1))
    grid = sdf.Grid(first, size, count)
    out = sdf.generate_grid_sdf_exact(v, topo, grid, sdf.SignMethod.Raycast)
    d2, _ = O.nearest_bvh(v, f, O.grid_centres(first, size, count))
    assert np.array_equal(bits(np.abs(out)), bits(np.sqrt(d2)))
    fill = sdf.generate_grid_sdf(v, topo, grid, sdf.SignMethod.Raycast)
    assert np.array_equal(np.signbit(out), np.signbit(fill))
    assert np.all(np.abs(out) <= np.abs(fill))
