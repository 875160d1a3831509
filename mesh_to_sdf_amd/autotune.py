"""Autotuner for libsdfx's launch-shape knobs (none changes a result).

    python -m mesh_to_sdf_amd.autotune [--out PATH]

measures, on the current GPU, each knob over its candidates on seeded
synthetic workloads and writes the fastest settings as JSON. `_lib` applies the
file `mesh_to_sdf_amd/tuning.json` (or the one named by SDFX_TUNING) when it
loads the library; SDFX_TUNING=none skips it.

Knobs:
* exhaustive_max_pairs: queries x triangles at or below which Bvh / Rtree /
  RtreeBvh answer with the exhaustive kernel instead of building an LBVH. The
  tuner times both implementations over a ladder of (queries, triangles) sizes
  and keeps the largest pair count at which the exhaustive one is still faster.
* grid_prop_blocks: blocks per grid propagation launch, timed on a 96^3 and a
  128^3 grid.
* grid_batch_prop_blocks: the same for batched grid calls, timed on the first
  32 requests of the batched workload.
* lbvh_visit_budget: node visits after which an LBVH nearest search moves to
  the per-wave pass. 0 (the library's default) picks 15 x cbrt(triangles) per
  mesh; the tuner times it against fixed budgets with Bvh(Raycast) on the
  4M-query noise-sphere workload (327,680 triangles, half the queries near the
  surface) and keeps 0 unless a fixed budget is more than 3% faster.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import time

import numpy as np

from . import _lib, synthetic as S

DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")


def _timed(fn, reps):
    import torch
    fn()  # warm-up
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best


def _device_point_call(L, sp, v, f, q):
    import torch
    dv = torch.from_numpy(v).cuda()
    df = torch.from_numpy(f.reshape(-1).astype(np.int32)).cuda()
    dq = torch.from_numpy(q).cuda()
    out = torch.empty(len(q), dtype=torch.float32, device="cuda")
    n = ctypes.c_uint64(0)

    def call():
        _lib.check(L.sdfx_generate_sdf_device(dv.data_ptr(), len(v), df.data_ptr(), df.numel(), 0, dq.data_ptr(),
                                              len(q), 3, 0, out.data_ptr(), ctypes.byref(n), sp))
    return call


def tune_exhaustive(L, sp, log):
    """Largest queries x triangles at which the exhaustive kernel beats the LBVH."""
    crossover = 0
    rows = []
    for level in (2, 3, 4, 5):
        v, f = S.icosphere(level)
        for nq in (1_000, 10_000, 100_000, 1_000_000):
            pairs = nq * len(f)
            if pairs > 1 << 33:  # keeps the exhaustive timings short
                continue
            q = S.uniform_queries(v, nq, seed=1000 + level)
            call = _device_point_call(L, sp, v, f, q)
            with _lib.tuned("exhaustive_max_pairs", 1 << 62):
                t_ex = _timed(call, 5)
            with _lib.tuned("exhaustive_max_pairs", 0):
                t_bvh = _timed(call, 5)
            rows.append({"triangles": len(f), "queries": nq, "pairs": pairs, "exhaustive_ms": t_ex * 1e3,
                         "lbvh_ms": t_bvh * 1e3})
            log(f"  {len(f):>7} tris x {nq:>8} queries: exhaustive {t_ex * 1e3:8.3f} ms, lbvh {t_bvh * 1e3:8.3f} ms")
    # the switch point: below the smallest size at which the LBVH wins
    for r in sorted(rows, key=lambda r: r["pairs"]):
        if r["lbvh_ms"] < r["exhaustive_ms"]:
            break
        crossover = r["pairs"]
    return crossover, rows


def tune_visit_budget(L, sp, log, candidates=(0, 384, 512, 768, 1024, 1536, 2048)):
    import torch
    v, f, q = S.config3()
    dv = torch.from_numpy(v).cuda()
    df = torch.from_numpy(f.reshape(-1).astype(np.int32)).cuda()
    dq = torch.from_numpy(q).cuda()
    out = torch.empty(len(q), dtype=torch.float32, device="cuda")
    n = ctypes.c_uint64(0)

    def call():  # Bvh(Raycast) through the LBVH
        _lib.check(L.sdfx_generate_sdf_device(dv.data_ptr(), len(v), df.data_ptr(), df.numel(), 0, dq.data_ptr(),
                                              len(q), 1, 0, out.data_ptr(), ctypes.byref(n), sp))
    ms = {}
    with _lib.tuned("exhaustive_max_pairs", 0):
        for c in candidates:
            with _lib.tuned("lbvh_visit_budget", c):
                ms[c] = _timed(call, 5) * 1e3
            log(f"  budget {c or 'by mesh size'}: {ms[c]:.3f} ms")
    best = min(ms, key=ms.get)
    return (0 if ms[0] <= 1.03 * ms[best] else best), {str(c): t for c, t in ms.items()}


def tune_grid_blocks(L, sp, log, candidates=(1024, 2048, 4096, 6144, 8192)):
    import torch
    v, f = S.torus()
    total = {c: 0.0 for c in candidates}
    for n in (96, 128):
        first, size, count = S.grid_for(v, (n, n, n))
        dv = torch.from_numpy(v).cuda()
        df = torch.from_numpy(f.reshape(-1).astype(np.int32)).cuda()
        out = torch.empty(n * n * n, dtype=torch.float32, device="cuda")
        fa = np.ascontiguousarray(first, np.float32)
        sa = np.ascontiguousarray(size, np.float32)
        ca = np.ascontiguousarray(count, np.uint64)
        F = ctypes.POINTER(ctypes.c_float)

        def call():
            _lib.check(L.sdfx_generate_grid_sdf_device(
                dv.data_ptr(), len(v), df.data_ptr(), df.numel(), 0, fa.ctypes.data_as(F), sa.ctypes.data_as(F),
                ca.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), 0, out.data_ptr(), sp))
        for c in candidates:
            with _lib.tuned("grid_prop_blocks", c):
                t = _timed(call, 10)
            total[c] += t
            log(f"  grid {n}^3, {c} blocks: {t * 1e3:.3f} ms")
    return min(total, key=total.get), {str(c): t * 1e3 for c, t in total.items()}


def tune_batch_grid_blocks(L, sp, log, candidates=(1024, 1536, 1792, 2048, 3072, 4096)):
    """grid_batch_prop_blocks, timed on the first 32 requests of the batched
    workload (one generate_grid_sdf_batch call each; the host copies are the
    same for every candidate)."""
    import mesh_to_sdf_amd as sdf
    reqs = [(v, sdf.Topology.TriangleList(f), sdf.Grid(first, size, count))
            for v, f, first, size, count in S.config4(n_meshes=32)]
    ms = {}
    for c in candidates:
        with _lib.tuned("grid_batch_prop_blocks", c):
            ms[c] = _timed(lambda: sdf.generate_grid_sdf_batch(reqs, sdf.SignMethod.Raycast), 3)
        log(f"  batch of {len(reqs)} grids, {c} blocks: {ms[c] * 1e3:.3f} ms")
    return min(ms, key=ms.get), {str(c): t * 1e3 for c, t in ms.items()}


def tune(out_path: str = DEFAULT_PATH, log=print) -> dict:
    import torch
    torch.cuda.set_device(0)
    L = _lib.lib()
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    log("exhaustive_max_pairs:")
    crossover, rows = tune_exhaustive(L, sp, log)
    log("lbvh_visit_budget:")
    budget, budget_ms = tune_visit_budget(L, sp, log)
    log("grid_prop_blocks:")
    blocks, grid_ms = tune_grid_blocks(L, sp, log)
    log("grid_batch_prop_blocks:")
    batch_blocks, batch_ms = tune_batch_grid_blocks(L, sp, log)
    result = {
        "device": torch.cuda.get_device_name(0),
        "knobs": {"exhaustive_max_pairs": int(crossover), "grid_prop_blocks": int(blocks),
                  "lbvh_visit_budget": int(budget), "grid_batch_prop_blocks": int(batch_blocks)},
        "measurements": {"exhaustive_vs_lbvh": rows, "lbvh_ms_by_visit_budget": budget_ms,
                         "grid_ms_by_blocks": grid_ms, "batch_grid_ms_by_blocks": batch_ms},
    }
    if out_path:
        with open(out_path, "w") as fh:
            json.dump(result, fh, indent=1)
        log(f"wrote {out_path}")
    return result


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_PATH)
    args = ap.parse_args()
    tune(args.out)


if __name__ == "__main__":
    main()
