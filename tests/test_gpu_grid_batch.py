"""Batched generate_grid_sdf: every request of a batch gets exactly the bytes
it gets alone (the schedule of each request reads only that request's state),
and a failing request does not disturb the others."""
import numpy as np
import pytest

from mesh_to_sdf_amd import synthetic as S

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def _requests(sdf, seed=3, n=12):
    g = np.random.default_rng(seed)
    reqs = []
    meshes = [S.icosphere(2), S.icosphere(3), S.torus(1.0, 0.35, 32, 16, 0.002), S.noise_sphere(3, 0.05),
              S.torus(1.0, 0.2, 48, 12, 0.001, seed=11)]
    for i in range(n):
        v, f = meshes[i % len(meshes)]
        counts = tuple(int(c) for c in g.integers(6, 40, size=3))
        first, size, count = S.grid_for(v, counts, factor=float(g.uniform(0.8, 1.3)))
        topo = sdf.Topology.TriangleList(f)
        reqs.append((v, topo, sdf.Grid(first, size, count)))
    return reqs


@pytest.mark.parametrize("sign", [0, 1])
def test_batch_equals_alone(sdf, sign):
    reqs = _requests(sdf)
    # extras: strip topology, a negative cell size, an empty mesh, a one-cell grid
    v, f = S.icosphere(2)
    strip = f.reshape(-1)[: 3 * 60].copy()
    reqs.append((v, sdf.Topology.TriangleStrip(strip), sdf.Grid(*S.grid_for(v, (9, 11, 13)))))
    reqs.append((v, sdf.Topology.TriangleList(f), sdf.Grid([1.2, -1.2, 1.2], [-0.1, 0.1, -0.1], [20, 20, 20])))
    reqs.append((np.zeros((0, 3), np.float32), sdf.Topology.TriangleList(np.zeros(0, np.uint32)),
                 sdf.Grid([0, 0, 0], [0.1, 0.1, 0.1], [4, 5, 6])))
    reqs.append((v, sdf.Topology.TriangleList(f), sdf.Grid([0.1, 0.2, 0.3], [0.5, 0.5, 0.5], [1, 1, 1])))
    batch = sdf.generate_grid_sdf_batch(reqs, sign)
    for i, (vv, topo, grid) in enumerate(reqs):
        alone = sdf.generate_grid_sdf(vv, topo, grid, sign)
        assert np.array_equal(bits(batch[i]), bits(alone)), f"request {i}"


def test_batch_matches_oracle(sdf, oracle):
    reqs = _requests(sdf, seed=5, n=4)
    batch = sdf.generate_grid_sdf_batch(reqs, 0)
    for i, (v, topo, grid) in enumerate(reqs):
        s, ref, _ = oracle.generate_grid_sdf(v, topo.indices, oracle.TOPO_LIST, grid.first_cell, grid.cell_size,
                                             grid.cell_count, 0, 8, oracle.GRID_FRONTIER)
        assert s == 0 and np.array_equal(bits(batch[i]), bits(ref)), f"request {i}"


def test_failing_request_is_isolated(sdf):
    from mesh_to_sdf_amd import _lib
    reqs = _requests(sdf, seed=9, n=5)
    v, f = S.icosphere(2)
    bad = f.reshape(-1).copy()
    bad[7] = len(v) + 3  # out of range
    reqs.insert(2, (v, sdf.Topology.TriangleList(bad), sdf.Grid(*S.grid_for(v, (10, 10, 10)))))
    reqs.insert(4, (v, sdf.Topology.TriangleList(f[:-1].reshape(-1)[:-1]), sdf.Grid(*S.grid_for(v, (8, 8, 8)))))
    reqs.append((v, sdf.Topology.TriangleList(f), sdf.Grid([0, 0, 0], [0.1, 0.1, 0.1], [5, 0, 5])))
    out, st = sdf.generate_grid_sdf_batch(reqs, 0, return_status=True)
    assert st[2] == _lib.ERR_INDEX_OUT_OF_RANGE
    assert st[4] == _lib.ERR_INDEX_COUNT
    assert st[-1] == _lib.ERR_INVALID_GRID
    for i, (vv, topo, grid) in enumerate(reqs):
        if i in (2, 4, len(reqs) - 1):
            assert out[i] is None
            continue
        assert st[i] == _lib.OK
        assert np.array_equal(bits(out[i]), bits(sdf.generate_grid_sdf(vv, topo, grid, 0))), f"request {i}"
    with pytest.raises(_lib.SdfxError) as e:
        sdf.generate_grid_sdf_batch(reqs, 0)
    assert e.value.status == _lib.ERR_INDEX_OUT_OF_RANGE and "request 2" in e.value.message


@pytest.mark.parametrize("sign", [0, 1])
def test_second_thin_round_release_matches_oracle(sdf, oracle, sign):
    """Grids whose result the release after a request's second thin round
    changes: alone and in one batch, bit-exact against the oracle's frontier
    schedule, which differs on them from the same schedule without the rule."""
    from test_grid_fixpoint import second_release_grids
    grids = list(second_release_grids())
    reqs = [(v, sdf.Topology.TriangleList(np.asarray(f, np.uint32).reshape(-1)), sdf.Grid(first, size, count))
            for v, f, (first, size, count) in grids]
    batch = sdf.generate_grid_sdf_batch(reqs, sign)
    differs = 0
    try:
        for i, (v, f, (first, size, count)) in enumerate(grids):
            f = np.asarray(f, np.uint32).reshape(-1)
            s, ref, _ = oracle.generate_grid_sdf(v, f, oracle.TOPO_LIST, first, size, count, sign, 8,
                                                 oracle.GRID_FRONTIER)
            assert s == 0
            alone = sdf.generate_grid_sdf(v, reqs[i][1], reqs[i][2], sign)
            assert np.array_equal(bits(alone), bits(ref)), f"grid {i} alone"
            assert np.array_equal(bits(batch[i]), bits(ref)), f"grid {i} in the batch"
            oracle.set_fault(oracle.FAULT_NO_SECOND_RELEASE)
            s, other, _ = oracle.generate_grid_sdf(v, f, oracle.TOPO_LIST, first, size, count, sign, 8,
                                                   oracle.GRID_FRONTIER)
            oracle.set_fault(oracle.FAULT_NONE)
            differs += not np.array_equal(bits(other), bits(ref))
    finally:
        oracle.set_fault(oracle.FAULT_NONE)
    assert differs >= 2
