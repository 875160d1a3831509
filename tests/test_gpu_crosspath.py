"""The reference's cross-path consistency and invariant tests, run on the
device paths (each path is also bit-exact against the oracle elsewhere; these
check the relations the reference itself asserts between them), plus
determinism across launch shapes."""
import numpy as np
import pytest

from mesh_to_sdf_amd import synthetic as S

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def _suzanne_grid(sdf, suzanne):
    v, f = suzanne
    grid = sdf.Grid.from_bounding_box(v.min(0), v.max(0), [32, 32, 32])
    ii = np.stack(np.meshgrid(*(np.arange(32),) * 3, indexing="ij"), -1).reshape(-1, 3)
    q = np.array([grid.get_cell_center(c) for c in ii], np.float32)
    return v, f.reshape(-1), grid, q


@pytest.mark.parametrize("method", ["Bvh", "RtreeBvh"])
def test_grid_vs_points_suzanne(sdf, suzanne, method):
    """generic/bvh.rs:287-295 and generic/rtree_bvh.rs:467-472: a 32^3 Raycast
    grid agrees with the point queries at the cell centres within 0.01."""
    v, f, grid, q = _suzanne_grid(sdf, suzanne)
    m = sdf.AccelerationMethod.Bvh(sdf.SignMethod.Raycast) if method == "Bvh" else sdf.AccelerationMethod.RtreeBvh
    pts = sdf.generate_sdf(v, sdf.Topology.TriangleList(f), q, m)
    g = sdf.generate_grid_sdf(v, sdf.Topology.TriangleList(f), grid, sdf.SignMethod.Raycast)
    assert np.abs(pts - g).max() < 0.01


def test_rtree_vs_grid_normal_suzanne(sdf, suzanne):
    """generic/rtree.rs:259-279: |d| within 0.01, fewer than 1% of signs differ."""
    v, f, grid, q = _suzanne_grid(sdf, suzanne)
    pts = sdf.generate_sdf(v, sdf.Topology.TriangleList(f), q, sdf.AccelerationMethod.Rtree)
    g = sdf.generate_grid_sdf(v, sdf.Topology.TriangleList(f), grid, sdf.SignMethod.Normal)
    assert np.abs(np.abs(pts) - np.abs(g)).max() < 0.01
    assert np.mean(np.abs(pts - g) > 0.01) < 0.01


def test_point_paths_agree(sdf, suzanne):
    """generic/bvh.rs:222-229 (Bvh vs None), generic/rtree.rs:196-203 (Rtree vs
    Bvh(Normal)), generic/rtree_bvh.rs:321-328 (RtreeBvh vs Bvh(Raycast)), each
    within 0.01: on the reference's five suzanne points; on 2,000 suzanne points
    for the magnitudes (suzanne is open, so a one-ray and a three-ray sign may
    differ away from those points); on 2,000 points around a watertight torus
    for the signed values."""
    M = sdf.AccelerationMethod
    R, N = sdf.SignMethod.Raycast, sdf.SignMethod.Normal
    pairs = [(M.Bvh(R), M.None_(R)), (M.Bvh(N), M.None_(N)), (M.Rtree, M.Bvh(N)), (M.RtreeBvh, M.Bvh(R))]
    v, f = suzanne
    T = sdf.Topology.TriangleList(f.reshape(-1))
    five = np.array([[0.01, 0.01, 0.5], [1.0, 1.0, 1.0], [0.1, 0.2, 0.2], [1.1, 2.2, 5.2], [-0.1, 0.2, -0.2]],
                    np.float32)
    q = np.concatenate([S.uniform_queries(v, 1000, 41), S.surface_queries(v, f.reshape(-1, 3), 1000, 0.02, 42)])
    tv, tf = S.torus(1.0, 0.35, 64, 32, 0.002, seed=8)
    tq = np.concatenate([S.uniform_queries(tv, 1000, 43), S.surface_queries(tv, tf, 1000, 0.02, 44)])
    for a, b in pairs:
        assert np.abs(sdf.generate_sdf(v, T, five, a) - sdf.generate_sdf(v, T, five, b)).max() < 0.01, (a, b)
        da, db = sdf.generate_sdf(v, T, q, a), sdf.generate_sdf(v, T, q, b)
        assert np.abs(np.abs(da) - np.abs(db)).max() < 0.01, (a, b)
        ta = sdf.generate_sdf(tv, sdf.Topology.TriangleList(tf), tq, a)
        tb = sdf.generate_sdf(tv, sdf.Topology.TriangleList(tf), tq, b)
        assert np.abs(ta - tb).max() < 0.01, (a, b)


def test_device_grid_continuity(sdf):
    """generate/grid.rs:765-803 on the device grid: neighbouring |d| differ by
    at most the cell size and signs flip only within a cell of the surface."""
    v, f = S.torus(1.0, 0.35, 48, 24, 0.0)
    lo, hi = S.aabb(v)
    ext = np.float32(0.2) * (hi - lo)
    grid = sdf.Grid.from_bounding_box(lo - ext, hi + ext, [32, 32, 32])
    g = sdf.generate_grid_sdf(v, sdf.Topology.TriangleList(f), grid, sdf.SignMethod.Raycast).reshape(32, 32, 32)
    size = grid.cell_size
    for axis in range(3):
        a = np.take(g, range(31), axis=axis)
        b = np.take(g, range(1, 32), axis=axis)
        assert np.all(np.abs(np.abs(a) - np.abs(b)) <= size[axis] * (1 + 1e-6))
        flip = (a * b) < 0
        assert np.all(np.abs(a[flip]) <= size[axis] * (1 + 1e-6))
        assert np.all(np.abs(b[flip]) <= size[axis] * (1 + 1e-6))


def test_results_independent_of_launch_shapes(sdf):
    """Every tuning knob changes launch shapes only: the grid, alone and
    batched, with different propagation grid sizes, and the point paths exhaustive or through the
    LBVH (with any visit budget for the per-query pass), give the same bits."""
    v, f, first, size, count = S.config2()
    grid = sdf.Grid(first, size, (64, 64, 64))
    T = sdf.Topology.TriangleList(f)
    outs = []
    for blocks in (8, 256, 4096, 9000):  # 8: each block strides over many pairs
        with sdf._lib.tuned("grid_prop_blocks", blocks):
            outs.append(sdf.generate_grid_sdf(v, T, grid, sdf.SignMethod.Raycast))
    assert all(np.array_equal(bits(outs[0]), bits(o)) for o in outs[1:])
    reqs = [(v, T, grid), (v, T, sdf.Grid(*S.grid_for(v, (24, 30, 18))))]
    for blocks in (8, 2048, 16384):
        with sdf._lib.tuned("grid_batch_prop_blocks", blocks):
            batch = sdf.generate_grid_sdf_batch(reqs, sdf.SignMethod.Raycast)
        assert np.array_equal(bits(batch[0]), bits(outs[0])), blocks
        assert np.array_equal(bits(batch[1]), bits(sdf.generate_grid_sdf(*reqs[1], sdf.SignMethod.Raycast))), blocks
    q = S.uniform_queries(v, 20000, 77)
    for m in (sdf.AccelerationMethod.Bvh(sdf.SignMethod.Raycast), sdf.AccelerationMethod.Rtree,
              sdf.AccelerationMethod.RtreeBvh):
        with sdf._lib.tuned("exhaustive_max_pairs", 0):
            a = sdf.generate_sdf(v, T, q, m)
            for budget in (16, 100, 2048, (1 << 31) - 1):  # per-wave pass for searches past the budget
                with sdf._lib.tuned("lbvh_visit_budget", budget):
                    assert np.array_equal(bits(a), bits(sdf.generate_sdf(v, T, q, m))), (m, budget)
        with sdf._lib.tuned("exhaustive_max_pairs", 1 << 62):
            b = sdf.generate_sdf(v, T, q, m)
        assert np.array_equal(bits(a), bits(b)), m
