"""The item-major propagation (k_frontier_propagate_items, for large rounds)
leaves exactly the keys the lane-per-neighbour kernel leaves: the
supporting-plane bound it applies skips only pairs that cannot improve their
neighbour. Forced for every round (grid_item_major_min = 0) and never used
(2^40), every grid here must come out bit for bit the same, and equal to the
oracle's restatement of the schedule."""
import numpy as np
import pytest

from mesh_to_sdf_amd import synthetic as S

from test_gpu_random_meshes import random_grid, soup

pytestmark = pytest.mark.gpu

ALWAYS, NEVER = 0, 1 << 40


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def both(sdf, fn):
    out = {}
    for knob in (ALWAYS, NEVER):
        with sdf._lib.tuned("grid_item_major_min", knob):
            out[knob] = fn()
    return out[ALWAYS], out[NEVER]


@pytest.mark.parametrize("sign", [0, 1])
def test_item_major_random_soups(sdf, oracle, sign):
    """Degenerate soups on random small grids (one-cell axes, negative sizes)."""
    for seed in range(7, 13):
        v, f, _ = soup(seed)
        g = np.random.default_rng(seed + 100)
        first, size, count = random_grid(g, v)
        topo = sdf.Topology.TriangleList(f.reshape(-1))
        a, b = both(sdf, lambda: sdf.generate_grid_sdf(v, topo, sdf.Grid(first, size, count), sign))
        s, ref, _ = oracle.generate_grid_sdf(v, f.reshape(-1), oracle.TOPO_LIST, first, size, count, sign, 4,
                                             oracle.GRID_FRONTIER)
        assert s == 0
        assert np.array_equal(bits(a), bits(b)), seed
        assert np.array_equal(bits(a), bits(ref)), seed


@pytest.mark.parametrize("sign", [0, 1])
def test_item_major_meshes_and_batch(sdf, oracle, suzanne, sign):
    """Suzanne at 40^3 (tie-prone in Normal mode) against the oracle, the
    bench grid, and a 24-request batch of the batched workload's meshes."""
    sv, sf = suzanne
    cases = [(np.asarray(sv, np.float32), np.asarray(sf, np.uint32).reshape(-1), S.grid_for(sv, (40, 40, 40)))]
    v, f, first, size, count = S.config2()
    cases.append((v, np.asarray(f, np.uint32).reshape(-1), (first, size, count)))
    for k, (v, f, (first, size, count)) in enumerate(cases):
        topo = sdf.Topology.TriangleList(f)
        a, b = both(sdf, lambda: sdf.generate_grid_sdf(v, topo, sdf.Grid(first, size, count), sign))
        assert np.array_equal(bits(a), bits(b)), k
        if k == 0:
            s, ref, _ = oracle.generate_grid_sdf(v, f, oracle.TOPO_LIST, first, size, count, sign, 4,
                                                 oracle.GRID_FRONTIER)
            assert np.array_equal(bits(a), bits(ref))
    reqs = [(mv, sdf.Topology.TriangleList(mf.reshape(-1)), sdf.Grid(mfirst, msize, mcount))
            for mv, mf, mfirst, msize, mcount in S.config4(24)]
    a, b = both(sdf, lambda: sdf.generate_grid_sdf_batch(reqs, sign))
    for k, (x, y) in enumerate(zip(a, b)):
        assert np.array_equal(bits(x), bits(y)), k
