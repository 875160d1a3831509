"""sdfx_query_sdf_batch_device: query_sdf (lib.rs:463-485) on many resident
meshes with uneven query counts in one call. Every answer equals query_sdf on
that mesh alone and the oracle's generate_sdf, bit for bit, for every
acceleration method; failing requests (an empty Bvh mesh, which query_sdf
answers with the reference's unwrap panic) leave the others intact."""
import numpy as np
import pytest

from mesh_to_sdf_amd import synthetic as S

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


@pytest.mark.parametrize("visit_budget", [0, 16])
def test_query_batch_equals_alone_and_oracle(sdf, oracle, visit_budget):
    meshes = [S.icosphere(2), S.torus(1.0, 0.35, 48, 24, 0.002, seed=5), S.noise_sphere(4, 0.05, seed=7),
              S.jittered_icosphere(3, 0.002, seed=9)]
    g = np.random.default_rng(21)
    accs, reqs, meta = [], [], []
    for k in range(30):
        mi = int(g.integers(len(meshes)))
        method = sdf.ALL_SDF_ACCELERATION_METHODS[k % 6]
        v, f = meshes[mi]
        acc = sdf.build_sdf_acceleration(v, sdf.Topology.TriangleList(f), method)
        n = int(np.exp(g.uniform(0, np.log(60_000)))) if k % 7 else 0
        q = np.zeros((0, 3), np.float32)
        if n:
            q = np.concatenate([S.uniform_queries(v, n - n // 2, 100 + k),
                                S.surface_queries(v, f, n // 2, 0.01 * S.bbox_diagonal(v), 200 + k)])
        accs.append(acc)
        reqs.append((acc, q))
        meta.append((mi, method))
    with sdf._lib.tuned("lbvh_visit_budget", visit_budget):
        outs = sdf.query_sdf_batch(reqs)
        for (acc, q), out, (mi, method) in zip(reqs, outs, meta):
            alone = sdf.query_sdf(acc, q)
            assert np.array_equal(bits(out), bits(alone)), (mi, method.kind, method.sign)
            v, f = meshes[mi]
            s, ref = oracle.generate_sdf(v, f, oracle.TOPO_LIST, q, method.kind, method.sign)
            assert np.array_equal(bits(out), bits(ref)), ("oracle", mi, method.kind)


def test_query_batch_failures_are_isolated(sdf):
    v, f = S.icosphere(2)
    good = sdf.build_sdf_acceleration(v, sdf.Topology.TriangleList(f), sdf.AccelerationMethod.Bvh(0))
    empty = sdf.build_sdf_acceleration(np.zeros((0, 3), np.float32), sdf.Topology.TriangleList(np.zeros(0, np.uint32)),
                                       sdf.AccelerationMethod.Bvh(0))
    q = S.uniform_queries(v, 777, 5)
    outs, st = sdf.query_sdf_batch([(good, q), (empty, q), (good, q[:10])], return_status=True)
    assert st[0] == 0 and st[2] == 0 and st[1] == sdf._lib.ERR_EMPTY_MESH and outs[1] is None
    assert np.array_equal(bits(outs[0]), bits(sdf.query_sdf(good, q)))
    assert np.array_equal(bits(outs[2]), bits(sdf.query_sdf(good, q[:10])))
