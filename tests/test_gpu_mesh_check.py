"""check_mesh_triangle_list (generate/mesh_check.rs:24-81) on the device
against the oracle's as-written restatement: the same verdict, the same
failing vertex / face / edge, the reference's message."""
import numpy as np
import pytest

import refcases as R
from mesh_to_sdf_amd import synthetic as S

pytestmark = pytest.mark.gpu

EX_V = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, -1.0]], np.float32)  # mesh_check.rs:88-89
EX_F = np.array([0, 1, 2, 0, 2, 3, 0, 3, 1, 1, 3, 2], np.uint32)


def expected_message(v, f, kind, where):
    f = np.asarray(f, np.uint32).reshape(-1)
    if kind == 0:
        return None
    if kind == 1:
        return "Faces must be a multiple of 3"
    if kind == 2:
        x, y, z = (R.rust_debug_f32(c) for c in np.asarray(v, np.float32).reshape(-1, 3)[where])
        return f"Duplicate vertex found at ({x}, {y}, {z})"
    if kind == 3:
        a, b, c = f[3 * where: 3 * where + 3]
        return f"Duplicate face found: [{a}, {b}, {c}]"
    face, k = divmod(where, 3)
    a, b = f[3 * face + k], f[3 * face + (k + 1) % 3]
    if kind == 4:
        return f"Duplicate edge {a} -> {b} - do you have an inverted normal?"
    return f"Mesh is not waterproof - edge {a} -> {b} does not have a corresponding edge {b} -> {a}"


def check_against_oracle(sdf, oracle, v, f, what):
    kind, where = oracle.check_mesh_triangle_list_at(v, f)
    got = sdf.check_mesh_triangle_list(v, f)
    assert got == expected_message(v, f, kind, where), f"{what}: oracle kind {kind} at {where}, device {got!r}"
    return kind


def test_reference_cases(sdf, oracle):
    """The five cases of mesh_check.rs:84-172."""
    assert check_against_oracle(sdf, oracle, EX_V, EX_F, "ok") == 0
    assert sdf.check_mesh_triangle_list(EX_V, EX_F) is None
    dup_v = np.concatenate([EX_V, EX_V[-1:]])
    assert check_against_oracle(sdf, oracle, dup_v, EX_F, "duplicate vertex") == 2
    assert sdf.check_mesh_triangle_list(dup_v, EX_F) == "Duplicate vertex found at (0.0, 0.0, -1.0)"
    dup_f = np.concatenate([EX_F, EX_F[-3:]])
    assert check_against_oracle(sdf, oracle, EX_V, dup_f, "duplicate face") == 3
    assert sdf.check_mesh_triangle_list(EX_V, dup_f) == "Duplicate face found: [1, 3, 2]"
    assert check_against_oracle(sdf, oracle, EX_V, EX_F[:-3], "missing face") == 5
    rev = EX_F.copy()
    rev[0], rev[2] = EX_F[2], EX_F[0]
    assert check_against_oracle(sdf, oracle, EX_V, rev, "reversed face") in (4, 5)
    assert check_against_oracle(sdf, oracle, EX_V, EX_F[:-1], "length") == 1


def test_vertex_identity_is_debug_text(sdf, oracle):
    """{:?} tells -0.0 from 0.0 and prints every NaN as "NaN"."""
    base = np.array([[0.0, 1.0, 2.0], [3.0, 4.0, 5.0], [6.0, 7.0, 8.0]], np.float32)
    f = np.array([0, 1, 2, 0, 2, 1], np.uint32)
    neg0 = np.concatenate([base, np.float32([[-0.0, 1.0, 2.0]])])
    assert check_against_oracle(sdf, oracle, neg0, f, "-0.0") != 2
    nan1 = np.float32([[np.nan, 1e-5, 3e38]])
    nan2 = nan1.copy()
    nan2.view(np.uint32)[0, 0] = 0xFFC01234  # another NaN payload, negative
    v = np.concatenate([base, nan1, nan2])
    assert check_against_oracle(sdf, oracle, v, f, "NaN payloads") == 2
    assert sdf.check_mesh_triangle_list(v, f) == "Duplicate vertex found at (NaN, 1e-5, 3e38)"


def _defective(g, v, f):
    f = f.reshape(-1, 3).copy()
    v = v.copy()
    kind = g.integers(0, 7)
    if kind == 0 and len(f) > 1:      # drop faces
        f = np.delete(f, g.choice(len(f), size=g.integers(1, 4), replace=False), axis=0)
    elif kind == 1:                   # flip faces
        idx = g.choice(len(f), size=g.integers(1, 4), replace=False)
        f[idx] = f[idx][:, ::-1]
    elif kind == 2:                   # duplicate a face
        f = np.concatenate([f, f[g.integers(0, len(f), size=1)]])
    elif kind == 3:                   # duplicate a vertex somewhere
        i, j = g.integers(0, len(v), size=2)
        v = np.insert(v, j, v[i], axis=0)
    elif kind == 4:                   # a face with a repeated index (loop edge)
        a, b = g.integers(0, len(v), size=2)
        f = np.insert(f, g.integers(0, len(f) + 1), [a, a, b], axis=0)
    elif kind == 5:                   # a stray face
        f = np.insert(f, g.integers(0, len(f) + 1), g.integers(0, len(v), size=3), axis=0)
    else:                             # rotate a face (same triangle, different start): duplicates stay unseen
        i = g.integers(0, len(f))
        f = np.concatenate([f, np.roll(f[i], 1)[None]])
    return v, f.reshape(-1)


@pytest.mark.parametrize("mesh", ["icosphere2", "torus16x8", "tetra"])
def test_random_defects_match_oracle(sdf, oracle, mesh):
    if mesh == "icosphere2":
        v, f = S.icosphere(2)
    elif mesh == "torus16x8":
        v, f = S.torus(1.0, 0.35, 16, 8, 0.0)
    else:
        v, f = EX_V, EX_F.reshape(-1, 3)
    g = np.random.default_rng(7)
    kinds = set()
    assert check_against_oracle(sdf, oracle, v, f.reshape(-1), f"{mesh} as built") == 0
    for case in range(60):
        dv, df = _defective(g, v, f.reshape(-1, 3))
        kinds.add(check_against_oracle(sdf, oracle, dv, df, f"{mesh} case {case}"))
    assert {2, 3, 5} <= kinds


def test_large_meshes(sdf):
    """The bench meshes are watertight; removing one face opens three edges."""
    v, f, *_ = S.config2()
    assert sdf.check_mesh_triangle_list(v, f) is None
    f2 = np.delete(f.reshape(-1, 3), 1234, axis=0)
    msg = sdf.check_mesh_triangle_list(v, f2)
    assert msg is not None and msg.startswith("Mesh is not waterproof - edge ")
    opened = {(int(x), int(y)) for x, y in zip(f.reshape(-1, 3)[1234], np.roll(f.reshape(-1, 3)[1234], -1))}
    edge = tuple(int(t) for t in msg.split("edge ")[1].split(" does")[0].split(" -> "))
    assert (edge[1], edge[0]) in opened  # an edge whose partner was in the removed face


def test_bench_meshes_watertight(sdf):
    """Every mesh the workloads run (the three single-mesh configurations, the
    256 of the batched workload and the 64 of the served workload) is
    watertight, the precondition of the Raycast sign (generate/mesh_check.rs)."""
    meshes = [S.config1()[:2], S.config2()[:2], S.config3(n_queries=16)[:2]]
    meshes += [(m[0], m[1]) for m in S.config4()]
    meshes += [(m[0], m[1]) for m in S.config4(64, seed=0x5DF00005)]  # the served workload's pool
    for i, (v, f) in enumerate(meshes):
        assert sdf.check_mesh_triangle_list(v, f) is None, f"mesh {i}"
