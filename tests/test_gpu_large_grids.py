"""Grids larger than one launch covers. A propagation launch takes at most
164.5M frontier items (its 32-bit work index covers 26 neighbour slots per
item), so larger rounds run several launches; a batch is cut into runs of at
most 2^28 tile-padded cells, each with its own scratch (the reference,
generate/grid.rs:265-378, has no cell limit)."""
import numpy as np
import pytest

from mesh_to_sdf_amd import synthetic as S

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def test_batch_over_one_run_equals_alone(sdf):
    """20 grids of 240^3-256^3 cells (325M cells in all, two runs) on two small
    meshes: every request equals the same request alone, bit for bit."""
    meshes = [S.icosphere(2), S.torus(1.0, 0.35, 24, 12, 0.002, seed=5)]
    reqs = []
    for k in range(20):
        v, f = meshes[k % 2]
        e = 256 - (k % 4) * 5
        reqs.append((v, sdf.Topology.TriangleList(f), sdf.Grid(*S.grid_for(v, (e, e, e), factor=1.1 + 0.01 * k))))
    total = sum(g.get_total_cell_count() for _, _, g in reqs)
    assert total > (1 << 28)
    outs = sdf.generate_grid_sdf_batch(reqs, sdf.SignMethod.Raycast)
    for k, ((v, topo, grid), out) in enumerate(zip(reqs, outs)):
        alone = sdf.generate_grid_sdf(v, topo, grid, sdf.SignMethod.Raycast)
        assert np.array_equal(bits(out), bits(alone)), k


def test_single_grid_over_one_propagation_launch(sdf, oracle):
    """A tilted quad through a slab of 600 x 600 x 500 cells (180M cells) whose
    two triangles' padded boxes hold every cell: every cell is seeded within
    one band, so the first round takes all 180M cells, beyond one launch. The
    seeds already hold the exact distances, so every sampled cell equals the
    brute-force FP32 distance; and the grid equals the same request inside a
    batch (the batched kernels), bit for bit."""
    v = np.array([[0, 0, 0], [600, 0, 0.5], [600, 600, 0.5], [0, 600, 0]], np.float32)
    f = np.array([0, 1, 2, 0, 2, 3], np.uint32)
    grid = sdf.Grid([0.5, 0.5, 0.0005], [1.0, 1.0, 0.001], [600, 600, 500])
    assert grid.get_total_cell_count() > 165_000_000
    topo = sdf.Topology.TriangleList(f)
    out = sdf.generate_grid_sdf(v, topo, grid, sdf.SignMethod.Raycast)
    g = np.random.default_rng(3)
    idx = np.sort(g.choice(out.size, 1_000_000, replace=False))
    z = idx % 500
    y = (idx // 500) % 600
    x = idx // (500 * 600)
    pts = np.stack([np.float32(0.5) + x.astype(np.float32) * np.float32(1.0),
                    np.float32(0.5) + y.astype(np.float32) * np.float32(1.0),
                    np.float32(0.0005) + z.astype(np.float32) * np.float32(0.001)], axis=1).astype(np.float32)
    d2, _ = oracle.nearest_bruteforce(v, f.reshape(-1, 3), pts)
    assert np.array_equal(bits(np.abs(out[idx])), bits(np.sqrt(d2)))
    small = (v, topo, sdf.Grid([0.5, 0.5, 0.0005], [10.0, 10.0, 0.01], [8, 9, 10]))
    batched = sdf.generate_grid_sdf_batch([(v, topo, grid), small], sdf.SignMethod.Raycast)
    assert np.array_equal(bits(batched[0]), bits(out))
