"""The served form (sdfx_server_*): each answer is bit-identical to the same
request answered alone through the plain entry points and to the oracle,
requests queued while the device is busy are served in shared batches, the
HBM budget gates meshes and requests, and a failing request leaves the others
untouched."""
import numpy as np
import pytest

from mesh_to_sdf_amd import synthetic as S

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def _meshes():
    return [S.icosphere(2), S.torus(1.0, 0.35, 24, 12, 0.002, seed=5), S.noise_sphere(3, 0.05, seed=7),
            S.jittered_icosphere(3, 0.002, seed=9)]


def _server(budget=4 << 30, workers=2):
    from mesh_to_sdf_amd.server import SdfServer
    return SdfServer(budget, workers)


def _oracle_answer(oracle, v, f, kind, a, b):
    """The oracle's answer to a served request: generate_sdf for points (method
    b), the frontier schedule for grids (sign b), which the plain grid path
    matches bit for bit."""
    if kind == "grid":
        s, ref, _ = oracle.generate_grid_sdf(v, f, oracle.TOPO_LIST, a.first_cell, a.cell_size, a.cell_count, int(b),
                                             16, oracle.GRID_FRONTIER)
    else:
        s, ref = oracle.generate_sdf(v, f, oracle.TOPO_LIST, a, b.kind, b.sign)
    assert s in (0, 1)
    return ref


@pytest.mark.parametrize("visit_budget", [1024, 16])
def test_mixed_requests_equal_alone(sdf, oracle, visit_budget):
    """Every method and both grid signs; with a 16-visit budget nearly every
    nearest search of the batched point pass finishes in the per-wave pass.
    Each answer equals the plain entry point's and the oracle's, bit for bit."""
    meshes = _meshes()
    g = np.random.default_rng(11)
    methods = list(sdf.ALL_SDF_ACCELERATION_METHODS)
    with sdf._lib.tuned("lbvh_visit_budget", visit_budget), _server() as srv:
        ids = [srv.add_mesh(v, sdf.Topology.TriangleList(f)) for v, f in meshes]
        jobs = []
        for k in range(36):
            mi = int(g.integers(len(meshes)))
            v, f = meshes[mi]
            if k % 3 == 2:
                counts = tuple(int(c) for c in g.integers(6, 48, size=3))
                grid = sdf.Grid(*S.grid_for(v, counts, factor=float(g.uniform(0.9, 1.3))))
                sign = int(g.integers(2))
                jobs.append(("grid", mi, grid, sign, srv.submit_grid(ids[mi], grid, sign)))
            else:
                n = int(g.integers(1, 6000))
                q = np.concatenate([S.uniform_queries(v, n - n // 2, seed=100 + k),
                                    S.surface_queries(v, f, n // 2, 0.01, seed=200 + k)])
                m = methods[k % len(methods)]
                jobs.append(("points", mi, q, m, srv.submit_points(ids[mi], q, m)))
        for kind, mi, a, b, tk in jobs:
            got, ms = srv.wait(tk)
            assert ms >= 0.0
            v, f = meshes[mi]
            if kind == "grid":
                alone = sdf.generate_grid_sdf(v, sdf.Topology.TriangleList(f), a, b)
            else:
                alone = sdf.generate_sdf(v, sdf.Topology.TriangleList(f), a, b)
            assert np.array_equal(bits(got.cpu().numpy()), bits(alone)), (kind, mi, b)
            ref = _oracle_answer(oracle, v, f, kind, a, b)
            assert np.array_equal(bits(got.cpu().numpy()), bits(ref)), ("oracle", kind, mi, b)
        st = srv.stats()
        assert st["requests_done"] == len(jobs) and st["queued"] == 0 and st["inflight_bytes"] == 0


def test_served_workload_alone_and_oracle(sdf, oracle):
    """A reduced served workload: 64 resident meshes drawn as in the batched
    workload (1,280-327,680 triangles, one batched admission), 32 requests
    drawn as the served stream draws them (points 1k-1M with Bvh(Raycast),
    grids 32^3-128^3 with Raycast) plus one 1M-point request and one 128^3 grid
    on the largest mesh, submitted in random order at Poisson arrivals (1,000
    requests/s) so that they join batches already running. Each answer equals
    the same request answered alone by the plain entry points, bit for bit;
    the requests of at most 100,000 points or 64^3 cells also equal the
    oracle's answer."""
    import time
    meshes = S.config4(64)
    reqs = S.serve_requests(meshes, 32, seed=S.SEED_BASE + 7)
    big = max(range(len(meshes)), key=lambda i: len(meshes[i][1]))
    vb, fb = meshes[big][0], meshes[big][1]
    reqs.append(("points", big, np.concatenate([S.surface_queries(vb, fb, 500_000, 0.01 * S.bbox_diagonal(vb), 31),
                                                S.uniform_queries(vb, 500_000, 32)])))
    reqs.append(("grid", big, S.grid_for(vb, (128, 128, 128))))
    method = sdf.AccelerationMethod.Bvh(sdf.SignMethod.Raycast)
    g = np.random.default_rng(6)
    with _server(budget=32 << 30) as srv:
        ids = srv.add_meshes([(v, sdf.Topology.TriangleList(f)) for v, f, *_ in meshes])
        jobs = []
        for k in g.permutation(len(reqs)):
            kind, mi, a = reqs[k]
            time.sleep(float(g.exponential(1e-3)))
            if kind == "points":
                jobs.append((kind, mi, a, srv.submit_points(ids[mi], a, method)))
            else:
                grid = sdf.Grid(*a)
                jobs.append((kind, mi, grid, srv.submit_grid(ids[mi], grid, sdf.SignMethod.Raycast)))
        checked = 0
        for kind, mi, a, tk in jobs:
            got, _ = srv.wait(tk)
            got = got.cpu().numpy()
            v, f = meshes[mi][0], meshes[mi][1]
            topo = sdf.Topology.TriangleList(f)
            if kind == "points":
                alone = sdf.generate_sdf(v, topo, a, method)
                small = len(a) <= 100_000
            else:
                alone = sdf.generate_grid_sdf(v, topo, a, sdf.SignMethod.Raycast)
                small = a.get_total_cell_count() <= 64 ** 3
            assert np.array_equal(bits(got), bits(alone)), (kind, mi)
            if small:
                ref = _oracle_answer(oracle, v, f, kind, a, method if kind == "points" else sdf.SignMethod.Raycast)
                assert np.array_equal(bits(got), bits(ref)), ("oracle", kind, mi)
                checked += 1
        st = srv.stats()
    print(f"served workload: {len(jobs)} requests in {st['batches']} batches (largest {st['max_batch_requests']}), "
          f"{checked} checked against the oracle")
    assert st["requests_done"] == len(jobs) and checked >= 8


def test_queued_requests_share_batches(sdf):
    """With one worker busy on a large grid, the requests submitted meanwhile
    wait in the queue and are then served together in one batch."""
    v, f = S.torus(1.0, 0.35, 128, 64, 0.002, seed=3)
    small = S.icosphere(2)
    with _server(workers=1) as srv:
        big = srv.add_mesh(v, sdf.Topology.TriangleList(f))
        sm = srv.add_mesh(small[0], sdf.Topology.TriangleList(small[1]))
        first = srv.submit_grid(big, sdf.Grid(*S.grid_for(v, (224, 224, 224))), 0)
        rest = []
        for k in range(12):
            q = S.uniform_queries(small[0], 500 + 37 * k, seed=k)
            rest.append((q, srv.submit_points(sm, q, sdf.AccelerationMethod.RtreeBvh)))
        srv.wait(first)
        for q, tk in rest:
            got, _ = srv.wait(tk)
            alone = sdf.generate_sdf(small[0], sdf.Topology.TriangleList(small[1]), q, sdf.AccelerationMethod.RtreeBvh)
            assert np.array_equal(bits(got.cpu().numpy()), bits(alone))
        st = srv.stats()
        assert st["max_batch_requests"] >= 2, st
        assert st["batches"] < 13, st


def test_budget_gates_meshes_and_requests(sdf):
    from mesh_to_sdf_amd import _lib
    v, f = S.icosphere(3)  # 1,280 triangles: 208 B each resident
    with _server(budget=2_000_000, workers=2) as srv:
        a = srv.add_mesh(v, sdf.Topology.TriangleList(f))
        big_v, big_f = S.icosphere(5)  # 20,480 triangles: 4.3 MB resident
        with pytest.raises(_lib.SdfxError) as e:
            srv.add_mesh(big_v, sdf.Topology.TriangleList(big_f))
        assert e.value.status == _lib.ERR_OVER_BUDGET
        with pytest.raises(_lib.SdfxError) as e:  # 128^3 cells need about 59 MB of scratch
            srv.submit_grid(a, sdf.Grid(*S.grid_for(v, (128, 128, 128))), 0)
        assert e.value.status == _lib.ERR_OVER_BUDGET
        # requests that fit one at a time but not all together still all run
        grids = [sdf.Grid(*S.grid_for(v, (30, 30, 30 + k))) for k in range(6)]  # about 1 MB of scratch each
        tks = [srv.submit_grid(a, gr, 0) for gr in grids]
        for gr, tk in zip(grids, tks):
            got, _ = srv.wait(tk)
            alone = sdf.generate_grid_sdf(v, sdf.Topology.TriangleList(f), gr, 0)
            assert np.array_equal(bits(got.cpu().numpy()), bits(alone))
        st = srv.stats()
        assert st["peak_inflight_bytes"] <= 2_000_000 - st["resident_bytes"], st


def test_failures_are_isolated(sdf):
    from mesh_to_sdf_amd import _lib
    v, f = S.icosphere(2)
    with _server() as srv:
        a = srv.add_mesh(v, sdf.Topology.TriangleList(f))
        empty = srv.add_mesh(np.zeros((0, 3), np.float32), sdf.Topology.TriangleList(np.zeros(0, np.uint32)))
        bad_idx = f.reshape(-1).copy()
        bad_idx[4] = len(v) + 1
        with pytest.raises(_lib.SdfxError) as e:
            srv.add_mesh(v, sdf.Topology.TriangleList(bad_idx))
        assert e.value.status == _lib.ERR_INDEX_OUT_OF_RANGE
        with pytest.raises(_lib.SdfxError) as e:
            srv.add_mesh(v, sdf.Topology.TriangleList(bad_idx[:-1]))
        assert e.value.status == _lib.ERR_INDEX_COUNT
        q = S.uniform_queries(v, 300, seed=1)
        t_ok = srv.submit_points(a, q, sdf.AccelerationMethod.Bvh(sdf.SignMethod.Raycast))
        t_zero = srv.submit_grid(a, sdf.Grid([0, 0, 0], [0.1, 0.1, 0.1], [4, 0, 4]), 0)
        t_empty_rb = srv.submit_points(empty, q, sdf.AccelerationMethod.RtreeBvh)
        t_empty_bvh = srv.submit_points(empty, q, sdf.AccelerationMethod.Bvh(sdf.SignMethod.Raycast))
        t_empty_grid = srv.submit_grid(empty, sdf.Grid([0, 0, 0], [0.1, 0.1, 0.1], [3, 4, 5]), 0)
        got, _ = srv.wait(t_ok)
        alone = sdf.generate_sdf(v, sdf.Topology.TriangleList(f), q, sdf.AccelerationMethod.Bvh(sdf.SignMethod.Raycast))
        assert np.array_equal(bits(got.cpu().numpy()), bits(alone))
        with pytest.raises(_lib.SdfxError) as e:
            srv.wait(t_zero)
        assert e.value.status == _lib.ERR_INVALID_GRID
        assert srv.wait(t_empty_rb)[0] is None  # rtree_bvh.rs:96-98: empty Vec
        with pytest.raises(_lib.SdfxError) as e:
            srv.wait(t_empty_bvh)
        assert e.value.status == _lib.ERR_EMPTY_MESH
        got, _ = srv.wait(t_empty_grid)  # no triangle: every cell keeps f32::MAX
        assert np.all(got.cpu().numpy() == np.finfo(np.float32).max)
        srv.remove_mesh(a)
        with pytest.raises(_lib.SdfxError) as e:
            srv.submit_points(a, q, sdf.AccelerationMethod.RtreeBvh)
        assert e.value.status == _lib.ERR_INVALID_ARGUMENT


def test_batched_mesh_admission(sdf):
    """Meshes added in one batched build answer exactly as meshes added one by
    one (and as the plain entry points); a failing mesh of the batch does not
    stop the others."""
    from mesh_to_sdf_amd import _lib
    meshes = _meshes() + [S.icosphere(1), (S.icosphere(0)[0], S.icosphere(0)[1][:1])]  # incl. one triangle
    g = np.random.default_rng(21)
    with _server() as one, _server() as many:
        ids1 = [one.add_meshes([(v, sdf.Topology.TriangleList(f))])[0] for v, f in meshes]
        ids2 = many.add_meshes([(v, sdf.Topology.TriangleList(f)) for v, f in meshes])
        for k, (v, f) in enumerate(meshes):
            q = S.uniform_queries(v, 3000, seed=300 + k)
            for m in (sdf.AccelerationMethod.Bvh(sdf.SignMethod.Raycast), sdf.AccelerationMethod.Bvh(sdf.SignMethod.Normal),
                      sdf.AccelerationMethod.Rtree, sdf.AccelerationMethod.RtreeBvh):
                a, _ = one.wait(one.submit_points(ids1[k], q, m))
                b, _ = many.wait(many.submit_points(ids2[k], q, m))
                ref = sdf.generate_sdf(v, sdf.Topology.TriangleList(f), q, m)
                assert np.array_equal(bits(a.cpu().numpy()), bits(b.cpu().numpy())), (k, m)
                assert np.array_equal(bits(b.cpu().numpy()), bits(ref)), (k, m)
        v, f = S.icosphere(2)
        bad = f.reshape(-1).copy()
        bad[3] = len(v) + 7
        descs = [(v, sdf.Topology.TriangleList(f)), (v, sdf.Topology.TriangleList(bad)),
                 (v, sdf.Topology.TriangleList(f))]
        with pytest.raises(_lib.SdfxError) as e:
            many.add_meshes(descs)
        assert e.value.status == _lib.ERR_INDEX_OUT_OF_RANGE and "mesh 1" in e.value.message
        assert many.stats()["meshes"] == len(meshes) + 2  # the two good ones of the batch were added


def test_cancel_queued_request(sdf):
    """A request still queued behind a busy worker can be cancelled; the others
    are unaffected, and a finished request cannot be."""
    from mesh_to_sdf_amd import _lib
    v, f = S.torus(1.0, 0.35, 128, 64, 0.002, seed=3)
    small = S.icosphere(2)
    with _server(workers=1) as srv:
        big = srv.add_mesh(v, sdf.Topology.TriangleList(f))
        sm = srv.add_mesh(small[0], sdf.Topology.TriangleList(small[1]))
        first = srv.submit_grid(big, sdf.Grid(*S.grid_for(v, (224, 224, 224))), 0)
        qs = [S.uniform_queries(small[0], 700, seed=50 + k) for k in range(3)]
        tks = [srv.submit_points(sm, q, sdf.AccelerationMethod.RtreeBvh) for q in qs]
        assert srv.cancel(tks[1])
        srv.wait(first)
        for k in (0, 2):
            got, _ = srv.wait(tks[k])
            alone = sdf.generate_sdf(small[0], sdf.Topology.TriangleList(small[1]), qs[k],
                                     sdf.AccelerationMethod.RtreeBvh)
            assert np.array_equal(bits(got.cpu().numpy()), bits(alone))
        with pytest.raises(_lib.SdfxError) as e:
            srv.wait(tks[1])
        assert e.value.status == _lib.ERR_CANCELLED
        t = srv.submit_points(sm, qs[0], sdf.AccelerationMethod.RtreeBvh)
        srv.wait(t)  # waited tickets are gone; a done, not yet waited one cannot be cancelled
        t2 = srv.submit_points(sm, qs[0], sdf.AccelerationMethod.RtreeBvh)
        while not srv.done(t2):
            pass
        assert not srv.cancel(t2)
        srv.wait(t2)


def test_concurrent_clients(sdf):
    """Several client threads submit and wait at once (the server's calls are
    thread-safe); every answer is the one the plain entry points give."""
    import threading
    meshes = _meshes()
    with _server(workers=3) as srv:
        ids = srv.add_meshes([(v, sdf.Topology.TriangleList(f)) for v, f in meshes])
        errors = []

        def client(c):
            try:
                g = np.random.default_rng(900 + c)
                for k in range(6):
                    mi = int(g.integers(len(meshes)))
                    v, f = meshes[mi]
                    if k % 2:
                        grid = sdf.Grid(*S.grid_for(v, tuple(int(x) for x in g.integers(8, 40, size=3))))
                        got, _ = srv.wait(srv.submit_grid(ids[mi], grid, 0))
                        ref = sdf.generate_grid_sdf(v, sdf.Topology.TriangleList(f), grid, 0)
                    else:
                        q = S.uniform_queries(v, int(g.integers(100, 5000)), seed=1000 * c + k)
                        got, _ = srv.wait(srv.submit_points(ids[mi], q, sdf.AccelerationMethod.RtreeBvh))
                        ref = sdf.generate_sdf(v, sdf.Topology.TriangleList(f), q, sdf.AccelerationMethod.RtreeBvh)
                    if not np.array_equal(bits(got.cpu().numpy()), bits(ref)):
                        errors.append((c, k))
            except Exception as e:  # reported below; a thread cannot fail the test itself
                errors.append((c, repr(e)))

        threads = [threading.Thread(target=client, args=(c,)) for c in range(6)]
        for t in threads:
            t.start()
        for t in threads:
            t.join(timeout=100)
        assert not any(t.is_alive() for t in threads)
        assert not errors, errors
        st = srv.stats()
        assert st["requests_done"] == 36 and st["queued"] == 0 and st["inflight_bytes"] == 0


def test_removed_mesh_stays_counted_until_its_requests_finish(sdf):
    """A mesh removed while a request for it waits behind a long grid keeps
    its bytes in the budget (its buffers are still in use) until that request
    is done; the request is answered as if the mesh were still there."""
    big_v, big_f = S.torus(1.0, 0.35, 128, 64, 0.002, seed=3)
    v, f = S.noise_sphere(4, 0.05, seed=7)
    q = S.uniform_queries(v, 50_000, 8)
    method = sdf.AccelerationMethod.Bvh(sdf.SignMethod.Raycast)
    with _server(workers=1) as srv:
        m0 = srv.add_mesh(big_v, sdf.Topology.TriangleList(big_f))
        r0 = srv.stats()["resident_bytes"]
        long_grid = srv.submit_grid(m0, sdf.Grid(*S.grid_for(big_v, (256, 256, 256))), 0)
        m1 = srv.add_mesh(v, sdf.Topology.TriangleList(f))
        r1 = srv.stats()["resident_bytes"]
        tk = srv.submit_points(m1, q, method)
        srv.remove_mesh(m1)
        st = srv.stats()
        assert st["meshes"] == 1 and st["resident_bytes"] == r1 > r0, st
        got, _ = srv.wait(tk)
        srv.wait(long_grid)
        assert srv.stats()["resident_bytes"] == r0
    alone = sdf.generate_sdf(v, sdf.Topology.TriangleList(f), q, method)
    assert np.array_equal(bits(got.cpu().numpy()), bits(alone))
