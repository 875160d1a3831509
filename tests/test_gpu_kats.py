"""The reference crate's known-answer tests, through the C ABI on the GPU."""
import math

import numpy as np
import pytest

import refcases as R

pytestmark = pytest.mark.gpu
f32 = np.float32


def test_tetrahedron_all_methods(sdf):
    """lib.rs:494-524 for ALL_SDF_ACCELERATION_METHODS."""
    topo = sdf.Topology.TriangleList(R.TETRA_INDICES)
    for m in sdf.ALL_SDF_ACCELERATION_METHODS:
        for q, expected in R.TETRA_POINTS:
            d = sdf.generate_sdf(R.TETRA_VERTICES, topo, [q], m)
            assert len(d) == 1
            assert float(d[0]) - expected < 1e-5, (m, q, d[0])
            if m != sdf.AccelerationMethod.Rtree:  # Rtree's sign comes from one of two tied faces
                assert abs(float(d[0]) - expected) < 1e-5, (m, q, d[0])


def test_open_and_reversed_tetrahedron(sdf):
    """lib.rs:527-555."""
    topo = sdf.Topology.TriangleList(R.TETRA_INDICES[:-3])
    d = sdf.generate_sdf(R.TETRA_VERTICES, topo, [[1.0, 1.0, -0.05]], sdf.AccelerationMethod.Bvh(sdf.SignMethod.Raycast))
    assert d[0] == np.sqrt(f32(0.05) * f32(0.05) + f32(0.5) * f32(0.5) + f32(0.5) * f32(0.5))
    idx = R.TETRA_INDICES.copy()
    idx[0], idx[2] = R.TETRA_INDICES[2], R.TETRA_INDICES[0]
    d = sdf.generate_sdf(R.TETRA_VERTICES, sdf.Topology.TriangleList(idx), [[0.2, 0.2, 0.05]],
                         sdf.AccelerationMethod.Bvh(sdf.SignMethod.Raycast))
    assert d[0] == f32(0.05)


def test_panics_become_errors(sdf):
    """lib.rs:557-585."""
    idx = R.TETRA_INDICES.copy()
    idx[0] = 100
    m = sdf.AccelerationMethod.Bvh(sdf.SignMethod.Raycast)
    with pytest.raises(sdf.SdfxError) as e:
        sdf.generate_sdf(R.TETRA_VERTICES, sdf.Topology.TriangleList(idx), [[0.2, 0.2, 0.05]], m)
    assert e.value.status == -4
    with pytest.raises(sdf.SdfxError) as e:
        sdf.generate_sdf(R.TETRA_VERTICES, sdf.Topology.TriangleList(np.append(R.TETRA_INDICES, 0)),
                         [[0.2, 0.2, 0.05]], m)
    assert e.value.status == -2 and "must be divisible by 3" in e.value.message
    with pytest.raises(sdf.SdfxError) as e:
        sdf.generate_sdf(R.TETRA_VERTICES, sdf.Topology.TriangleList(R.TETRA_INDICES), [[0.0, 0.0, np.nan]], m)
    assert e.value.status == -1 and "non-finite" in e.value.message


def test_cube(sdf):
    """lib.rs:588-683."""
    topo = sdf.Topology.TriangleList(R.CUBE_INDICES)
    for m in (sdf.AccelerationMethod.None_(sdf.SignMethod.Normal), sdf.AccelerationMethod.Bvh(sdf.SignMethod.Normal),
              sdf.AccelerationMethod.Rtree):
        d = sdf.generate_sdf(R.CUBE_VERTICES, topo, R.CUBE_INSIDE, m)
        assert np.all(np.abs(d - f32(-0.2)) < 1e-6), (m, d)
        d = sdf.generate_sdf(R.CUBE_VERTICES, topo, R.CUBE_OUTSIDE, m)
        assert np.all(np.abs(d - R.CUBE_OUTSIDE_EXPECTED) < 1e-6), (m, d)


def test_closest_vectors(sdf):
    """generic/rtree_bvh.rs:383-419."""
    acc = sdf.build_sdf_acceleration(R.TETRA_VERTICES, sdf.Topology.TriangleList(R.TETRA_INDICES),
                                     sdf.AccelerationMethod.RtreeBvh)
    assert acc is not None
    q = np.array([p for p, _, _ in R.CLOSEST_VECTOR_POINTS], np.float32)
    vec, inside = sdf.query_vector_to_closest_point_rtree_bvh(acc, q)
    for k, (p, ev, ein) in enumerate(R.CLOSEST_VECTOR_POINTS):
        assert np.linalg.norm(vec[k] - np.array(ev, np.float32)) < 1e-6, (p, vec[k])
        if np.linalg.norm(vec[k]) > 1e-6:
            assert bool(inside[k]) == ein


def test_doctests(sdf):
    """lib.rs:31, lib.rs:58, lib.rs:362, lib.rs:461, generate/grid.rs:231."""
    topo = sdf.Topology.TriangleList(np.array([0, 1, 2], np.uint32))
    assert list(sdf.generate_sdf(R.DOC_VERTICES_1, topo, R.DOC_QUERY_1, sdf.AccelerationMethod.RtreeBvh)) == [1.0]
    assert list(sdf.generate_sdf(R.DOC_VERTICES_2, topo, R.DOC_QUERY_2, sdf.AccelerationMethod.RtreeBvh)) == [1.0]
    acc = sdf.build_sdf_acceleration(R.DOC_VERTICES_2, topo, sdf.AccelerationMethod.RtreeBvh)
    assert list(sdf.query_sdf(acc, R.DOC_QUERY_2)) == [1.0]
    grid = sdf.Grid.from_bounding_box([0, 0, 0], [10, 10, 10], [10, 10, 10])
    for v in (R.DOC_VERTICES_1, R.DOC_VERTICES_GRID):
        g = sdf.generate_grid_sdf(v, topo, grid, sdf.SignMethod.Raycast)
        assert g[0] == 1.0


def test_grid_equals_none_raycast(sdf):
    """generate/grid.rs:693-724: exact equality on the 5^3 grid."""
    grid = sdf.Grid.from_bounding_box([0, 0, 0], [5, 5, 5], [5, 5, 5])
    topo = sdf.Topology.TriangleList(R.GRID2_INDICES)
    q = grid.cells_centers()
    d = sdf.generate_sdf(R.GRID2_VERTICES, topo, q, sdf.AccelerationMethod.None_(sdf.SignMethod.Raycast))
    g = sdf.generate_grid_sdf(R.GRID2_VERTICES, topo, grid, sdf.SignMethod.Raycast)
    assert np.array_equal(d, g)


def test_suzanne_cross_method(sdf, suzanne):
    """generic/bvh.rs:342-388, generic/rtree.rs:166-210, generic/rtree_bvh.rs:570-616,
    default.rs:116-151: cross-method tolerance 0.01, pysdf baseline 0.1, and
    query_sdf(build(..)) identical to generate_sdf."""
    v, idx = suzanne
    topo = sdf.Topology.TriangleList(idx)
    M = sdf.AccelerationMethod
    q = R.SUZANNE_POINTS
    bvh_r = sdf.generate_sdf(v, topo, q, M.Bvh(sdf.SignMethod.Raycast))
    none_r = sdf.generate_sdf(v, topo, q, M.None_(sdf.SignMethod.Raycast))
    assert np.all(np.abs(bvh_r - none_r) < 0.01)
    rtree = sdf.generate_sdf(v, topo, q, M.Rtree)
    bvh_n = sdf.generate_sdf(v, topo, q, M.Bvh(sdf.SignMethod.Normal))
    assert np.all(np.abs(rtree - bvh_n) < 0.01)
    rb = sdf.generate_sdf(v, topo, q, M.RtreeBvh)
    assert np.all(np.abs(rb - bvh_r) < 0.01)
    base = sdf.generate_sdf(v, topo, R.SUZANNE_BASELINE_POINTS, M.None_(sdf.SignMethod.Normal))
    assert np.all(np.abs(base - np.array(R.SUZANNE_BASELINE, np.float32)) < 0.1)
    for m in sdf.ALL_SDF_ACCELERATION_METHODS:
        acc = sdf.build_sdf_acceleration(v, topo, m)
        a = sdf.query_sdf(acc, q)
        b = sdf.generate_sdf(v, topo, q, m)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), m


def test_empty_mesh_conventions(sdf):
    """None: f32::MAX per query; RtreeBvh: empty Vec and build() -> None; Rtree: panic."""
    v = np.zeros((0, 3), np.float32)
    topo = sdf.Topology.TriangleList(np.zeros(0, np.uint32))
    q = [[0.0, 0.0, 0.0], [1.0, 2.0, 3.0]]
    d = sdf.generate_sdf(v, topo, q, sdf.AccelerationMethod.None_(sdf.SignMethod.Raycast))
    assert list(d) == [np.finfo(np.float32).max] * 2
    assert len(sdf.generate_sdf(v, topo, q, sdf.AccelerationMethod.RtreeBvh)) == 0
    assert sdf.build_sdf_acceleration(v, topo, sdf.AccelerationMethod.RtreeBvh) is None
    with pytest.raises(sdf.SdfxError) as e:
        sdf.generate_sdf(v, topo, q, sdf.AccelerationMethod.Rtree)
    assert e.value.status == -5
    grid = sdf.Grid.from_bounding_box([0, 0, 0], [1, 1, 1], [4, 4, 4])
    g = sdf.generate_grid_sdf(v, topo, grid, sdf.SignMethod.Raycast)
    assert np.all(g == np.finfo(np.float32).max)
