"""The grid's propagation pinned by the state every schedule of the reference
reaches, not by a restatement of one schedule.

The reference's flood fill (generate/grid.rs:495-558) pops states from a heap
(one per rayon thread, generate/grid.rs:318-339) and offers each popped
state's triangle to the cell's 27 neighbours, storing it where
compare_distances (lib.rs:315-332) calls it Less. Whatever the thread count
or timing, when a cell's last pushed state is popped, its triangle t(c) is
offered to every neighbour, and values only ever move earlier in
compare_distances' order afterwards. So at the end, for every cell c:

  (1) value(c) is the distance from c's centre to t(c), bit for bit;
  (2) no neighbour n has d(centre(n), t(c)) Less than value(n).

The oracle's frontier schedule (the device's schedule, which the GPU tests
match bit for bit) and its restatements of the reference's 1-heap and 8-heap
schedules must all satisfy this, with zero violations, and a schedule with a
planted mistake must not. On top, accuracy against the exact FP32 distance
(brute-force minimum over all triangles) is compared with the reference's
1-heap schedule: the frontier schedule must be at least as close.
"""
import itertools

import numpy as np
import pytest

from mesh_to_sdf_amd import synthetic as S

import refcases as R


def _fix(O, v, f, first, size, count, sign, out, tri):
    return O.grid_fixpoint_check(v, f, O.TOPO_LIST, first, size, count, sign, out, tri)


CLEAN = {"value_mismatch": 0, "not_fixpoint": 0, "no_triangle": 0, "no_triangle_value": 0}


def test_compare_distances_is_a_strict_weak_order(oracle):
    """The device keeps each cell's minimum under compare_distances with a
    compare-and-swap loop, which gives an order-independent result only if
    compare_distances is transitive. Checked on clusters of near-tied values
    of both signs (within a few ulps and within the 1e-6 epsilon) at
    magnitudes from 1e-8 to 100, every ordered triple."""
    O = oracle
    rng = np.random.default_rng(11)
    triples = 0
    for _ in range(600):
        base = np.float32(10.0 ** rng.uniform(-8, 2))
        vals = []
        for _ in range(4):
            mode = rng.integers(3)
            if mode == 0:
                v = np.uint32(int(base.view(np.uint32)) + int(rng.integers(-20, 21))).view(np.float32)
            elif mode == 1:
                v = np.float32(base + rng.uniform(-3e-6, 3e-6))
            else:
                v = base
            v = np.float32(abs(float(v)))
            vals.append(-v if rng.integers(2) else v)
        for a, b, c in itertools.permutations(vals, 3):
            triples += 1
            ab, bc, ac = O.compare_distances(a, b), O.compare_distances(b, c), O.compare_distances(a, c)
            assert ab == -O.compare_distances(b, a), (a, b)
            if ab == -1 and bc == -1:
                assert ac == -1, (a, b, c)
            if ab == 0 and bc == 0:
                assert ac == 0, (a, b, c)
    assert triples == 600 * 24


def test_config2_every_schedule_reaches_the_fixpoint(oracle):
    """The bench grid (jittered torus, 65,536 triangles, 128^3 cells, Raycast):
    the frontier schedule and the reference's 1-heap schedule both reach the
    fixpoint with zero violations, no value is below the exact FP32 distance,
    and the frontier schedule is at least as accurate as the reference's
    1-heap run (cells beyond 1e-5 x bbox diagonal, and the maximum error)."""
    O = oracle
    v, f, first, size, count = S.config2()
    diag = S.bbox_diagonal(v)
    d2, _ = O.nearest_bvh(v, f, O.grid_centres(first, size, count))
    exact = np.sqrt(d2)
    stats = {}
    for name, heaps, sched in (("frontier", 8, O.GRID_FRONTIER), ("reference 1-heap", 1, O.GRID_SEQUENTIAL)):
        s, out, _, tri = O.generate_grid_sdf(v, f, O.TOPO_LIST, first, size, count, O.SIGN_RAYCAST, heaps, sched,
                                             want_tri=True)
        assert s == 0
        assert _fix(O, v, f, first, size, count, O.SIGN_RAYCAST, out, tri) == CLEAN, name
        err = (np.abs(out) - exact) / diag
        assert err.min() >= 0.0, name
        stats[name] = (int(np.sum(err > 1e-5)), float(err.max()), float(np.mean(np.abs(out) != exact)))
        print(f"{name}: {stats[name][0]} cells beyond 1e-5 x diag, max error {stats[name][1]:.3g} x diag, "
              f"{stats[name][2]:.2%} off the exact distance")
    assert stats["frontier"][0] <= stats["reference 1-heap"][0]
    assert stats["frontier"][1] <= stats["reference 1-heap"][1]


@pytest.mark.parametrize("sign", [0, 1])
def test_small_grids_every_schedule_reaches_the_fixpoint(oracle, sign):
    """Both signs on a torus, a noise sphere and the reference's tetrahedron,
    grids with uneven axes; 1-heap, 8-heap and frontier schedules."""
    O = oracle
    meshes = [S.torus(1.0, 0.35, 48, 24, 0.002), S.noise_sphere(3, 0.05), (R.TETRA_VERTICES, R.TETRA_INDICES)]
    for v, f in meshes:
        v = np.asarray(v, np.float32)
        first, size, count = S.grid_for(v, (21, 17, 24))
        for heaps, sched in ((1, O.GRID_SEQUENTIAL), (8, O.GRID_SEQUENTIAL), (8, O.GRID_FRONTIER)):
            s, out, _, tri = O.generate_grid_sdf(v, f, O.TOPO_LIST, first, size, count, sign, heaps, sched,
                                                 want_tri=True)
            assert s == 0
            assert _fix(O, v, f, first, size, count, sign, out, tri) == CLEAN, (len(f), heaps, sched)


def second_release_grids():
    """Grids whose result the release after a second thin round changes (the
    band minimum keeps falling below the far-field distance after a first
    release, fronts meeting in a torus' hole): two tori and two requests of
    the batched workload."""
    v, f = S.torus(n_major=128, n_minor=64)
    yield v, f, S.grid_for(v, (80, 80, 80))
    v, f = S.torus(n_major=64, n_minor=32)
    yield v, f, S.grid_for(v, (96, 96, 96))
    c4 = S.config4(24)
    for i in (13, 23):
        mv, mf, first, size, count = c4[i]
        yield mv, mf.reshape(-1), (first, size, count)


@pytest.mark.parametrize("sign", [0, 1])
def test_second_thin_round_release(oracle, sign):
    """The release after a request's second thin round takes fewer rounds on
    these grids and changes the frontier schedule's result on at least two of
    them (the rule is exercised), and both schedules, with and without it,
    reach the reference's fixpoint."""
    O = oracle
    changed = 0
    try:
        for v, f, (first, size, count) in second_release_grids():
            runs = {}
            for fault in (O.FAULT_NONE, O.FAULT_NO_SECOND_RELEASE):
                O.set_fault(fault)
                s, out, st, tri = O.generate_grid_sdf(v, f, O.TOPO_LIST, first, size, count, sign, 8,
                                                      O.GRID_FRONTIER, want_tri=True)
                O.set_fault(O.FAULT_NONE)
                assert s == 0
                assert _fix(O, v, f, first, size, count, sign, out, tri) == CLEAN, (count, fault)
                runs[fault] = (out, st.rounds)
            assert runs[O.FAULT_NONE][1] < runs[O.FAULT_NO_SECOND_RELEASE][1], count
            changed += not np.array_equal(runs[O.FAULT_NONE][0], runs[O.FAULT_NO_SECOND_RELEASE][0])
    finally:
        O.set_fault(O.FAULT_NONE)
    assert changed >= 2


def test_fixpoint_check_catches_planted_schedule_faults(oracle, suzanne):
    """A frontier schedule that never offers a triangle to the +x neighbour, or
    that drops the cells beyond a round's band, leaves violations; so does
    ordering Normal-mode values by the packed key alone (|d| first), which
    keeps a negative value where compare_distances prefers an approximately
    equal positive one."""
    O = oracle
    v, f = S.torus(1.0, 0.35, 64, 32, 0.002)
    first, size, count = S.grid_for(v, (40, 40, 40))
    try:
        for fault in (O.FAULT_SKIP_DIRECTION, O.FAULT_DROP_DEFERRED):
            O.set_fault(fault)
            s, out, _, tri = O.generate_grid_sdf(v, f, O.TOPO_LIST, first, size, count, O.SIGN_RAYCAST, 8,
                                                 O.GRID_FRONTIER, want_tri=True)
            O.set_fault(O.FAULT_NONE)
            c = _fix(O, v, f, first, size, count, O.SIGN_RAYCAST, out, tri)
            assert c["not_fixpoint"] > 0, (fault, c)
        sv, sf = suzanne
        first, size, count = S.grid_for(sv, (40, 40, 40))
        O.set_fault(O.FAULT_PACKED_ORDER)
        s, out, _, tri = O.generate_grid_sdf(sv, sf, O.TOPO_LIST, first, size, count, O.SIGN_NORMAL, 8,
                                             O.GRID_FRONTIER, want_tri=True)
        O.set_fault(O.FAULT_NONE)
        c = _fix(O, sv, sf, first, size, count, O.SIGN_NORMAL, out, tri)
        assert c["not_fixpoint"] > 0, c
    finally:
        O.set_fault(O.FAULT_NONE)


@pytest.mark.parametrize("mesh,n", [("suzanne", 25), ("suzanne", 32), ("suzanne", 40), ("cube", 25), ("cube", 40)])
def test_normal_mode_ties_follow_compare_distances(oracle, suzanne, mesh, n):
    """Normal sign on tie-prone meshes (shared edges and vertices): every
    schedule reaches the fixpoint under compare_distances, and the frontier
    schedule's signs equal the reference 1-heap schedule's except at a few
    cells where the frontier holds either exactly the same |d| from another
    triangle, positive (the value compare_distances prefers; the 1-heap
    schedule never offered that triangle to the cell), or a strictly smaller
    |d| that is still at least the exact distance (a nearer triangle the 1-heap
    schedule never offered, whose sign differs)."""
    O = oracle
    if mesh == "suzanne":
        v, f = suzanne
    else:
        v, f = R.CUBE_VERTICES, R.CUBE_INDICES
    v = np.asarray(v, np.float32)
    first, size, count = S.grid_for(v, (n, n, n))
    outs = {}
    for heaps, sched in ((1, O.GRID_SEQUENTIAL), (8, O.GRID_SEQUENTIAL), (8, O.GRID_FRONTIER)):
        s, out, _, tri = O.generate_grid_sdf(v, f, O.TOPO_LIST, first, size, count, O.SIGN_NORMAL, heaps, sched,
                                             want_tri=True)
        assert s == 0
        assert _fix(O, v, f, first, size, count, O.SIGN_NORMAL, out, tri) == CLEAN
        outs[(heaps, sched)] = out
    fr, h1 = outs[(8, O.GRID_FRONTIER)], outs[(1, O.GRID_SEQUENTIAL)]
    bad = np.nonzero(np.signbit(fr) != np.signbit(h1))[0]
    print(f"{mesh} {n}^3 Normal: {bad.size} sign mismatches frontier vs 1-heap")
    assert bad.size <= 4
    if bad.size:
        exact = np.sqrt(O.nearest_bvh(v, f, O.grid_centres(first, size, count)[bad])[0])
        afr, ah1 = np.abs(fr[bad]), np.abs(h1[bad])
        tie = afr == ah1
        assert np.all(fr[bad][tie] > 0)
        assert np.all(afr[~tie] < ah1[~tie]) and np.all(afr >= exact)


def _slivers(seed, n=200):
    """Nearly collinear triangles (third vertex within 1e-7..1e-2 of the line
    through the first two, inside or past the segment) offset up to 10^4 from
    the origin, with a random grid around them."""
    g = np.random.default_rng(seed)
    off = np.float32(10.0 ** g.uniform(0, 4))
    a = g.uniform(-1, 1, (n, 3))
    d = g.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    length = g.uniform(0.01, 0.5, (n, 1))
    b = a + d * length
    eps = 10.0 ** g.uniform(-7, -2, (n, 1))
    c = a + d * length * g.uniform(-0.5, 1.5, (n, 1)) + g.normal(size=(n, 3)) * eps * length
    v = (np.concatenate([a, b, c]) + off).astype(np.float32)
    f = np.stack([np.arange(n), np.arange(n) + n, np.arange(n) + 2 * n], 1).astype(np.uint32)
    lo, hi = v.min(0), v.max(0)
    cnt = [int(x) for x in g.integers(16, 40, 3)]
    size = ((hi - lo) * 1.2 / np.array(cnt)).astype(np.float32)
    first = (lo - 0.1 * (hi - lo) + size / 2).astype(np.float32)
    return v, f.reshape(-1), first, size, cnt


def _bound_cases(suzanne):
    from test_gpu_random_meshes import random_grid, soup
    sv, sf = suzanne
    yield "torus", *S.torus(1.0, 0.35, 48, 24, 0.002), (24, 20, 28), 0
    yield "noise sphere", *S.noise_sphere(3, 0.05), (30, 30, 30), 1
    yield "suzanne", np.asarray(sv, np.float32), np.asarray(sf, np.uint32), (32, 32, 32), 1
    for seed in range(12):
        v, f, _ = soup(seed)
        first, size, count = random_grid(np.random.default_rng(seed + 500), v)
        yield f"soup {seed}", v, f.reshape(-1), (first, size, [min(48, 2 * c) for c in count]), seed & 1
    for seed in range(12):
        v, f, first, size, count = _slivers(seed)
        yield f"slivers {seed}", v, f, (first, size, count), seed & 1


def _run_bound(O, v, f, grid, sign, mode):
    if len(grid) == 3 and np.ndim(grid[0]) == 0:
        first, size, count = S.grid_for(v, grid)
    else:
        first, size, count = grid
    O.set_bound_check(mode)
    try:
        s, _, _ = O.generate_grid_sdf(v, f, O.TOPO_LIST, first, size, count, sign, 4, O.GRID_FRONTIER)
        assert s == 0
        return O.bound_check_counts()
    finally:
        O.set_bound_check(O.BOUND_OFF)


@pytest.mark.parametrize("mode", [1, 2])
def test_plane_bound_skips_only_pairs_that_cannot_improve(oracle, suzanne, mode):
    """The device propagation skips a (neighbour, triangle) pair when the
    supporting-plane bound g + u.o already exceeds the neighbour's key
    distance (plane_bound in sdfx_grid.hip, restated in FP32 by the oracle).
    Every such pair of the frontier schedule is evaluated anyway and must not
    lower its neighbour's key: with the direction words improvements write
    (mode 1) and with arbitrary ones (mode 2, the bound holds for any unit
    vector), on meshes, degenerate soups and slivers up to 10^4 from the
    origin, both signs. With exact directions the bound skips at least a
    third of the pairs it tests on the meshes."""
    O = oracle
    total = {"checked": 0, "skipped": 0, "wrong": 0}
    for name, v, f, grid, sign in _bound_cases(suzanne):
        c = _run_bound(O, v, f, grid, sign, mode)
        assert c["wrong"] == 0, (name, c)
        for k in total:
            total[k] += c[k]
        if mode == 1 and name in ("torus", "noise sphere", "suzanne"):
            assert c["skipped"] >= c["checked"] // 3, (name, c)
    print(f"mode {mode}: {total}")
    assert total["skipped"] > 0


def test_plane_bound_check_catches_a_wrong_margin(oracle, suzanne):
    """A bound whose rounding margin is added instead of subtracted skips
    pairs that would improve: the check sees it on the same cases."""
    O = oracle
    wrong = 0
    O.set_fault(O.FAULT_BOUND_MARGIN_SIGN)
    try:
        for name, v, f, grid, sign in _bound_cases(suzanne):
            wrong += _run_bound(O, v, f, grid, sign, O.BOUND_EXACT_DIRECTION)["wrong"]
    finally:
        O.set_fault(O.FAULT_NONE)
    print(f"wrong margin sign: {wrong} skipped pairs that improve")
    assert wrong > 0
