"""Accuracy of the grid schedules against the exact FP32 distance (run as a
script: python tests/grid_accuracy_table.py; CPU only, a few minutes).

For each grid: the oracle's frontier schedule (the device's, bit for bit) and
the reference's heap schedule (generate/grid.rs:318-339) with 1, 8 and 16
heaps run one after another (one legal order of that many rayon threads).
Each entry is: cells farther than 1e-5 x bbox diagonal from the exact
distance / largest error over the diagonal / fraction of cells whose value is
not the exact distance. The exact distance of a cell centre is the FP32
minimum over all triangles of geo.rs's distance (the oracle's BVH query,
equal to brute force)."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from mesh_to_sdf_amd import synthetic as S  # noqa: E402


def grids():
    v, f, first, size, count = S.config2()
    yield "bench grid: torus 65,536 tris, 128^3", v, f, (first, size, count)
    yield "torus 65,536 tris, 96^3", v, f, S.grid_for(v, (96, 96, 96))
    v, f = S.torus(n_major=128, n_minor=64, seed=S.SEED_BASE + 20)
    yield "torus 16,384 tris, 64^3", v, f, S.grid_for(v, (64, 64, 64))
    v, f = S.noise_sphere(5, 0.05, seed=S.SEED_BASE + 21)
    yield "noise sphere 20,480 tris, 96x64x80", v, f, S.grid_for(v, (96, 64, 80))
    v, f = S.jittered_icosphere(5, seed=S.SEED_BASE + 22)
    yield "icosphere 20,480 tris, 48^3", v, f, S.grid_for(v, (48, 48, 48))
    v, f = S.jittered_icosphere(3, seed=S.SEED_BASE + 23)
    yield "icosphere 1,280 tris, 64^3", v, f, S.grid_for(v, (64, 64, 64))
    yield "icosphere 1,280 tris, 37^3", v, f, S.grid_for(v, (37, 37, 37))


def entry(out, exact, diag):
    err = (np.abs(out) - exact) / diag
    assert err.min() >= 0.0
    return f"{int(np.sum(err > 1e-5)):,} / {err.max():.2e} / {np.mean(np.abs(out) != exact):.2%}"


def main():
    print("| grid | device (frontier rounds) | reference, 1 heap | reference, 8 heaps | reference, 16 heaps |")
    print("|---|---|---|---|---|")
    for name, v, f, (first, size, count) in grids():
        f = np.asarray(f, np.uint32).reshape(-1)
        diag = S.bbox_diagonal(v)
        exact = np.sqrt(O.nearest_bvh(v, f.reshape(-1, 3), O.grid_centres(first, size, count))[0])
        row = [name]
        for heaps, sched in ((8, O.GRID_FRONTIER), (1, O.GRID_SEQUENTIAL), (8, O.GRID_SEQUENTIAL),
                             (16, O.GRID_SEQUENTIAL)):
            s, out, _ = O.generate_grid_sdf(v, f, O.TOPO_LIST, first, size, count, O.SIGN_RAYCAST, heaps, sched)[:3]
            assert s == 0
            row.append(entry(out, exact, diag))
        print("| " + " | ".join(row) + " |", flush=True)


if __name__ == "__main__":
    main()
