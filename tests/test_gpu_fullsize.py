"""The benchmark workloads at their full sizes, checked through properties that
do not need the oracle to run at that size:

* the noise-sphere point workload (327,680 triangles, 4,194,304 queries): the
  LBVH path (per-query pass, then the per-wave pass for searches past the visit
  budget) gives the same bits as the exhaustive kernel, an independent
  algorithm that evaluates every triangle for every query, and every answer is
  finite;
* the batched grid workload (256 meshes, 84.3M cells): every request of one
  batched call gets the same bits as the same request alone;
* a single 17.3M-cell grid: the same bits alone and inside a batch.
The oracle pins both paths at smaller sizes (test_gpu_parity.py,
test_gpu_grid_batch.py)."""
import numpy as np
import pytest

from mesh_to_sdf_amd import synthetic as S

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


@pytest.mark.parametrize("method", ["Bvh", "Rtree"])
def test_noise_sphere_4m_lbvh_equals_exhaustive(sdf, method):
    v, f, q = S.config3()
    T = sdf.Topology.TriangleList(f)
    m = sdf.AccelerationMethod.Bvh(sdf.SignMethod.Raycast) if method == "Bvh" else sdf.AccelerationMethod.Rtree
    with sdf._lib.tuned("exhaustive_max_pairs", 0):
        lbvh = sdf.generate_sdf(v, T, q, m)
    with sdf._lib.tuned("exhaustive_max_pairs", 1 << 62):
        exhaustive = sdf.generate_sdf(v, T, q, m)
    assert np.all(np.isfinite(lbvh))
    bad = np.flatnonzero(bits(lbvh) != bits(exhaustive))
    assert bad.size == 0, f"{bad.size} of {len(q)} differ, first {bad[:5]}"


def test_batched_grids_256_equal_alone(sdf):
    reqs = [(v, sdf.Topology.TriangleList(f), sdf.Grid(first, size, count))
            for v, f, first, size, count in S.config4()]
    batch = sdf.generate_grid_sdf_batch(reqs, sdf.SignMethod.Raycast)
    for i, (v, topo, grid) in enumerate(reqs):
        alone = sdf.generate_grid_sdf(v, topo, grid, sdf.SignMethod.Raycast)
        assert np.array_equal(bits(batch[i]), bits(alone)), f"request {i}"


def test_noise_sphere_4m_closest_vectors_finite_and_consistent(sdf):
    """query_vector_to_closest_point on the 4M-query workload: every vector is
    finite and its length is the RtreeBvh distance's magnitude (both come from
    the same nearest triangle), with and without the per-wave pass."""
    v, f, q = S.config3()
    T = sdf.Topology.TriangleList(f)
    acc = sdf.build_sdf_acceleration(v, T, sdf.AccelerationMethod.RtreeBvh)
    vec, inside = sdf.query_vector_to_closest_point_rtree_bvh(acc, q)
    with sdf._lib.tuned("lbvh_visit_budget", (1 << 31) - 1):
        vec1, inside1 = sdf.query_vector_to_closest_point_rtree_bvh(acc, q)
    assert np.array_equal(bits(vec), bits(vec1)) and np.array_equal(inside, inside1)
    assert np.all(np.isfinite(vec))
    d = sdf.query_sdf(acc, q)
    assert np.array_equal(inside.astype(bool), d < 0)
    assert np.allclose(np.linalg.norm(vec.astype(np.float64), axis=1), np.abs(d), rtol=1e-5, atol=1e-7)
    acc.close()


def test_large_single_grid_equals_batched(sdf):
    """A single grid above 4,096 tiles of 4,096 cells (264 x 256 x 256 = 17.3M
    cells), where the compaction reads cell flags only for flagged tiles: the
    same bits as the request inside a batch (next to a small second request),
    whose kernels address requests through the batch table instead."""
    v, f, _, _, _ = S.config2()
    first, size, count = S.grid_for(v, (264, 256, 256))
    assert count[0] * count[1] * count[2] > 4096 * 4096
    T = sdf.Topology.TriangleList(f)
    grid = sdf.Grid(first, size, count)
    alone = sdf.generate_grid_sdf(v, T, grid, sdf.SignMethod.Raycast)
    small = sdf.Grid(*S.grid_for(v, (20, 20, 20)))
    batch = sdf.generate_grid_sdf_batch([(v, T, grid), (v, T, small)], sdf.SignMethod.Raycast)
    assert np.all(np.isfinite(alone))
    assert np.array_equal(bits(batch[0]), bits(alone))
    assert np.array_equal(bits(batch[1]), bits(sdf.generate_grid_sdf(v, T, small, sdf.SignMethod.Raycast)))
