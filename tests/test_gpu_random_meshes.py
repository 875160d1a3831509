"""Random triangle soups with the degenerate cases the reference's geometry
distinguishes (geo.rs:73-88: a point, a segment a-c, a segment a-b; shared and
duplicated vertices), queried at points where distances tie (vertices, edge
midpoints, face centroids, points on the padded boxes' faces) as well as at
random points: every method, through the LBVH (with its usual visit budget and
with nearly every search finished by the per-wave pass) and through the
exhaustive kernels, bit for bit against the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

METHODS = [(0, 0), (0, 1), (1, 0), (1, 1), (2, 1), (3, 0)]
PATHS = {"lbvh": {"exhaustive_max_pairs": 0, "lbvh_visit_budget": 0},
         "lbvh_deferred": {"exhaustive_max_pairs": 0, "lbvh_visit_budget": 16},
         "exhaustive": {"exhaustive_max_pairs": 1 << 62, "lbvh_visit_budget": 0}}


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def soup(seed, n_tris=300):
    g = np.random.default_rng(seed)
    # a small vertex pool so vertices are shared, with exact duplicates
    pool = g.uniform(-1, 1, size=(n_tris // 2, 3)).astype(np.float32)
    pool = np.concatenate([pool, pool[: len(pool) // 8]])  # duplicated positions
    f = g.integers(0, len(pool), size=(n_tris, 3)).astype(np.uint32)
    k = n_tris // 10
    f[:k, 1] = f[:k, 0]                    # a == b: segment a-c
    f[k:2 * k, 2] = f[k:2 * k, 1]          # b == c: segment a-b
    f[2 * k:2 * k + 3, 1:] = f[2 * k:2 * k + 3, :1]  # a == b == c: a point
    v = pool
    tri = v[f]
    q = [g.uniform(-1.3, 1.3, size=(400, 3)),
         v[g.integers(0, len(v), size=100)],                       # on vertices
         0.5 * (tri[:100, 0] + tri[:100, 1]),                      # on edges
         tri[100:200].mean(axis=1),                                # on faces
         tri[200:260].min(axis=1) - np.float32(0.0001)]            # on padded box corners
    return v, f, np.concatenate(q).astype(np.float32)


def _accel(sdf, accel, sign):
    M = sdf.AccelerationMethod
    return [M.None_(sign), M.Bvh(sign), M.Rtree, M.RtreeBvh][accel]


@pytest.mark.parametrize("path", sorted(PATHS))
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_soups_all_methods(sdf, oracle, seed, path):
    v, f, q = soup(seed)
    T = sdf.Topology.TriangleList(f.reshape(-1))
    with sdf._lib.tuned("exhaustive_max_pairs", PATHS[path]["exhaustive_max_pairs"]), \
            sdf._lib.tuned("lbvh_visit_budget", PATHS[path]["lbvh_visit_budget"]):
        for accel, sign in METHODS:
            s, ref = oracle.generate_sdf(v, f.reshape(-1), oracle.TOPO_LIST, q, accel, sign)
            assert s == 0
            gpu = sdf.generate_sdf(v, T, q, _accel(sdf, accel, sign))
            bad = np.flatnonzero(bits(gpu) != bits(ref))
            assert bad.size == 0, (seed, path, accel, sign, bad[:5], gpu[bad[:5]], ref[bad[:5]])


def random_grid(g, v):
    lo, hi = v.min(0), v.max(0)
    count = [int(c) for c in g.integers(1, 24, size=3)]
    scale = np.float32(g.uniform(1.0, 1.4))
    size = ((hi - lo) * scale / np.array(count, np.float32)).astype(np.float32)
    first = (lo - size * np.float32(0.5)).astype(np.float32)
    flip = g.integers(0, 4)
    if flip < 3:  # a negative cell size on one axis: the grid runs backwards from its first cell
        first[flip] = hi[flip] + size[flip] * np.float32(0.5)
        size[flip] = -size[flip]
    return first, size, count


@pytest.mark.parametrize("sign", [0, 1])
@pytest.mark.parametrize("seed", [4, 5, 6])
def test_random_soups_grids(sdf, oracle, seed, sign):
    """Degenerate soups on random small grids (one-cell axes, negative cell
    sizes), alone and batched, against the oracle's restatement of the device
    schedule."""
    g = np.random.default_rng(seed)
    reqs, refs = [], []
    for k in range(4):
        v, f, _ = soup(seed * 10 + k, n_tris=120)
        first, size, count = random_grid(g, v)
        s, ref, _ = oracle.generate_grid_sdf(v, f.reshape(-1), oracle.TOPO_LIST, first, size, count, sign, 8,
                                             oracle.GRID_FRONTIER)
        assert s == 0
        reqs.append((v, sdf.Topology.TriangleList(f.reshape(-1)), sdf.Grid(first, size, count)))
        refs.append(ref)
    batch = sdf.generate_grid_sdf_batch(reqs, sign)
    for k, ((v, topo, grid), ref) in enumerate(zip(reqs, refs)):
        alone = sdf.generate_grid_sdf(v, topo, grid, sign)
        assert np.array_equal(bits(alone), bits(ref)), (seed, sign, k, "alone")
        assert np.array_equal(bits(batch[k]), bits(ref)), (seed, sign, k, "batched")
