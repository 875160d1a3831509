"""SDF samples/s of the MI355X path on one GPU.

Default workload (the metric's served form): a server (sdfx_server_*) holding
64 resident meshes drawn as in the batched workload (jittered icospheres and
tori of 1,280-327,680 triangles, admitted with one batched LBVH build) answers
256 requests drawn as synthetic.serve_requests draws them: half point queries
(1k-1M points, log-uniform, half near-surface) with Bvh(Raycast), half grids
(32^3-128^3, log-uniform edge) with Raycast. One step = the 256 requests
submitted at once, until the last is answered (saturated throughput); one
SDF sample = one output float (a point's distance or a grid cell). Also
reported under `forms`: Poisson arrivals at each rate in SERVE_LAMBDAS
(samples/s, p50/p99 latency from submission to completion), and the plain
drop-in (sdfx_generate_sdf_device / sdfx_generate_grid_sdf_device, which
rebuild their acceleration per call) running the same requests one by one.
Inputs are resident in HBM before the timed region.

    python bench.py [--steps K] [--warmup W] [--workload serve|grid|grid_batch|rtree_bvh|bvh|rtree|none|mesh_check]

grid: one step = one generate_grid_sdf (Raycast) on a jittered torus (R=1,
r=0.35, 256x128 segments, 65,536 triangles) over a 128^3 grid from its AABB
expanded by 10%. grid_batch: one step = one sdfx_generate_grid_sdf_batch_device
call over 256 requests, reported beside the plain loop of single calls.
rtree_bvh / bvh / rtree / none: one generate_sdf call per step.

Prints one JSON line. The CPU baseline is the oracle's restatement of the
reference's threaded algorithm, timed on this host's cores (the CPU share the
machine gives one GPU) on a bounded sample; with it comes an accuracy block:
sign mismatches against the oracle with the samples within 1e-5 x bbox
diagonal of the surface excluded (and counted), and the error against an FP64
brute-force distance for the GPU and the CPU alike.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOADS = {
    "grid": "generate_grid_sdf Raycast, jittered torus 65,536 tris, 128^3 cells",
    "rtree_bvh": "generate_sdf RtreeBvh, icosphere L3 1,280 tris, 10,000 uniform queries",
    "none": "generate_sdf None(Raycast), icosphere L3 1,280 tris, 1,048,576 uniform queries",
    "rtree": "generate_sdf Rtree, noise sphere L7 327,680 tris, 4,194,304 queries",
    "bvh": "generate_sdf Bvh(Raycast), noise sphere L7 327,680 tris, 4,194,304 queries",
    "grid_batch": "batched generate_grid_sdf Raycast, 256 jittered icospheres/tori (1,280-327,680 tris, "
                  "log-uniform), one grid each with axes log-uniform in [32, 128]",
    "mesh_check": "check_mesh_triangle_list on the noise sphere L7 (327,680 tris, 163,842 vertices)",
    "serve": "served requests over 64 resident meshes (as in the batched workload): 256 requests, half point "
             "queries (1k-1M points, log-uniform, half near-surface) with Bvh(Raycast), half grids (edge "
             "log-uniform in [32, 128]) with Raycast",
}
SERVE_LAMBDAS = (50, 100, 200, 500, 1000, 2000)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="serve", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-kernel-timing", action="store_true",
                    help="skip the HIP-event kernel-timing pass (for runs under an external profiler)")
    ap.add_argument("--burst-only", action="store_true",
                    help="serve: time the burst steps only (no Poisson sweep or plain run), for profiler passes "
                         "that must see only the timed region's launches")
    ap.add_argument("--serve-workers", type=int, default=2,
                    help="serve: worker threads of the server (each with its own stream)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run three 256 MiB device-to-device copies (the reference for scaling the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    return ap.parse_args()


def calibrate_copy():
    """Three copies of 256 MiB (read 256 MiB, write 256 MiB each): under
    `rocprofv3 --pmc FETCH_SIZE` (or WRITE_SIZE) they give the counter's reading
    for a known number of HBM bytes."""
    import torch
    n = 256 * 1024 * 1024 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copies": 3, "bytes_read_per_copy": 4 * n, "bytes_written_per_copy": 4 * n}))


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` in `workload` from the committed counter
    summary (profiles/traffic.json, keyed "workload/kernel": mean FETCH_SIZE /
    WRITE_SIZE KiB per dispatch, each scaled by the reading the same counter
    gives for a copy of known size), or None when there is no entry for it."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        t = json.load(fh).get(f"{workload}/{kernel}")
    if not t:
        return None
    return 1024.0 * (t["fetch_kib_per_dispatch"] * t["fetch_scale"] + t["write_kib_per_dispatch"] * t["write_scale"])


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: every GPU machine here has one device; "
                                   "run with --gpus 1"}))
        sys.exit(2)
    import numpy as np
    import torch

    import mesh_to_sdf_amd as m
    from mesh_to_sdf_amd import _lib, synthetic as S

    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")
    L = _lib.lib()
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)

    if args.workload == "grid":
        v, f, first, size, count = S.config2()
        dv = torch.from_numpy(v).to(dev)
        df = torch.from_numpy(f.reshape(-1).astype(np.int32)).to(dev)
        ncells = count[0] * count[1] * count[2]
        dout = torch.empty(ncells, dtype=torch.float32, device=dev)
        first_a = np.ascontiguousarray(first, np.float32)
        size_a = np.ascontiguousarray(size, np.float32)
        cnt_a = np.ascontiguousarray(count, np.uint64)
        F = ctypes.POINTER(ctypes.c_float)

        def step():
            _lib.check(L.sdfx_generate_grid_sdf_device(
                ctypes.c_void_p(dv.data_ptr()), len(v), ctypes.c_void_p(df.data_ptr()), df.numel(), 0,
                first_a.ctypes.data_as(F), size_a.ctypes.data_as(F), cnt_a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                0, ctypes.c_void_p(dout.data_ptr()), sp))

        samples = ncells
        mesh_bytes = 12 * len(v) + 12 * len(f)
        config = {"workload": WORKLOADS["grid"], "triangles": int(len(f)), "cells": int(ncells),
                  "sign": "Raycast"}
    elif args.workload == "grid_batch":
        meshes = S.config4()
        reqs = (_lib.GridRequest * len(meshes))()
        resident = []  # device buffers kept alive for the run
        samples = tri_total = mesh_bytes = 0
        for r, (mv, mf, mfirst, msize, mcount) in zip(reqs, meshes):
            dvi = torch.from_numpy(mv).to(dev)
            dfi = torch.from_numpy(mf.reshape(-1).astype(np.int32)).to(dev)
            nc = mcount[0] * mcount[1] * mcount[2]
            doi = torch.empty(nc, dtype=torch.float32, device=dev)
            resident.append((dvi, dfi, doi))
            r.vertices, r.vertex_count = dvi.data_ptr(), len(mv)
            r.indices, r.index_count = dfi.data_ptr(), dfi.numel()
            r.topology = 0
            r.first_cell[:] = [float(x) for x in mfirst]
            r.cell_size[:] = [float(x) for x in msize]
            r.cell_count[:] = list(mcount)
            r.out = doi.data_ptr()
            samples += nc
            tri_total += len(mf)
            mesh_bytes += 12 * len(mv) + 12 * len(mf)
        v, f = meshes[0][0], meshes[0][1]
        F = ctypes.POINTER(ctypes.c_float)
        U = ctypes.POINTER(ctypes.c_uint64)

        def step():
            _lib.check(L.sdfx_generate_grid_sdf_batch_device(reqs, len(meshes), 0, sp))

        def plain_call(i):
            r = reqs[i]
            _lib.check(L.sdfx_generate_grid_sdf_device(
                ctypes.c_void_p(r.vertices), r.vertex_count, ctypes.c_void_p(r.indices), r.index_count, 0,
                ctypes.cast(r.first_cell, F), ctypes.cast(r.cell_size, F), ctypes.cast(r.cell_count, U), 0,
                ctypes.c_void_p(r.out), sp))

        config = {"workload": WORKLOADS["grid_batch"], "requests": len(meshes), "triangles": int(tri_total),
                  "cells": int(samples), "sign": "Raycast"}
    elif args.workload == "mesh_check":
        v, f, _ = S.config3(n_queries=16)
        fv = np.ascontiguousarray(f.reshape(-1), np.uint32)
        kind, where = ctypes.c_int(0), ctypes.c_uint64(0)
        msg = ctypes.create_string_buffer(512)
        Fp = ctypes.POINTER(ctypes.c_float)

        def step():
            st = L.sdfx_check_mesh_triangle_list(v.ctypes.data_as(Fp), len(v), fv.ctypes.data_as(
                ctypes.POINTER(ctypes.c_uint32)), fv.size, ctypes.byref(kind), ctypes.byref(where), msg, 512)
            if st != _lib.OK:
                raise RuntimeError(f"bench mesh is not watertight: {msg.value.decode()}")

        samples = len(f)
        mesh_bytes = 12 * len(v) + 12 * len(f)
        config = {"workload": WORKLOADS["mesh_check"], "triangles": int(len(f)), "vertices": int(len(v)),
                  "unit_of_work": "triangles checked (host arrays in, verdict out)"}
    elif args.workload == "serve":
        serve = ServeWorkload(torch, dev, L, sp, workers=args.serve_workers)
        step = serve.burst
        samples, mesh_bytes = serve.samples, serve.mesh_bytes
        v, f = serve.meshes[0][0], serve.meshes[0][1]
        config = {"workload": WORKLOADS["serve"], "meshes": len(serve.meshes), "requests": len(serve.reqs),
                  "point_queries": serve.point_samples, "grid_cells": serve.grid_samples,
                  "workers": serve.workers, "hbm_budget_bytes": serve.budget}
    else:
        if args.workload == "rtree_bvh":
            v, f, q = S.config1()
            method = (3, 0)
        elif args.workload == "none":
            v, f = S.icosphere(3)
            q = S.uniform_queries(v, 1 << 20, S.SEED_BASE + 11)
            method = (0, 0)
        else:
            v, f, q = S.config3()
            method = (2, 1) if args.workload == "rtree" else (1, 0)
        dv = torch.from_numpy(v).to(dev)
        df = torch.from_numpy(f.reshape(-1).astype(np.int32)).to(dev)
        dq = torch.from_numpy(q).to(dev)
        dout = torch.empty(len(q), dtype=torch.float32, device=dev)
        n = ctypes.c_uint64(0)

        def step():
            _lib.check(L.sdfx_generate_sdf_device(
                ctypes.c_void_p(dv.data_ptr()), len(v), ctypes.c_void_p(df.data_ptr()), df.numel(), 0,
                ctypes.c_void_p(dq.data_ptr()), len(q), method[0], method[1], ctypes.c_void_p(dout.data_ptr()),
                ctypes.byref(n), sp))

        samples = len(q)
        mesh_bytes = 12 * len(v) + 12 * len(f)
        config = {"workload": WORKLOADS[args.workload], "triangles": int(len(f)), "queries": int(len(q))}

    for _ in range(args.warmup):
        step()

    def timed(k):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            step()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    step_ms = []
    torch.cuda.synchronize()
    t_all = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        step()
        torch.cuda.synchronize()
        step_ms.append((time.perf_counter() - t0) * 1e3)
    elapsed = time.perf_counter() - t_all
    ms = elapsed * 1000.0 / args.steps
    value = samples * args.steps / elapsed
    forms = None
    if args.workload == "grid_batch":
        # The same requests as a plain loop of single-grid calls: each call
        # returns when its grid is done, so its time is that request's latency.
        for i in range(len(meshes)):
            plain_call(i)
        lat, loops = [], max(1, min(3, args.steps))
        torch.cuda.synchronize()
        t_loop = time.perf_counter()
        for _ in range(loops):
            for i in range(len(meshes)):
                t0 = time.perf_counter()
                plain_call(i)
                torch.cuda.synchronize()
                lat.append((time.perf_counter() - t0) * 1e3)
        loop_s = (time.perf_counter() - t_loop) / loops
        pct = lambda a, p: float(np.percentile(np.asarray(a), p))
        forms = {"batched": {"samples_per_s": value, "ms_per_batch": ms, "p50_ms": pct(step_ms, 50),
                             "p99_ms": pct(step_ms, 99),
                             "latency": "every request of a batch completes with the batch call"},
                 "plain_loop": {"samples_per_s": samples / loop_s, "ms_per_pass": loop_s * 1e3,
                                "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
                                "latency": "per request (one sdfx_generate_grid_sdf_device call)"},
                 "batched_over_plain": value / (samples / loop_s)}
    if args.workload == "serve" and not args.burst_only:
        forms = serve.forms(value, ms, step_ms)
    if args.workload == "grid":
        # what a drop-in caller pays: the host-pointer entry point, which also
        # uploads the mesh and downloads the distances over PCIe (not `value`)
        hv = np.ascontiguousarray(v, np.float32)
        hf = np.ascontiguousarray(f.reshape(-1), np.uint32)
        hout = np.empty(ncells, np.float32)
        U32 = ctypes.POINTER(ctypes.c_uint32)

        def host_call():
            _lib.check(L.sdfx_generate_grid_sdf(
                hv.ctypes.data_as(F), len(v), hf.ctypes.data_as(U32), hf.size, 0, first_a.ctypes.data_as(F),
                size_a.ctypes.data_as(F), cnt_a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), 0,
                hout.ctypes.data_as(F)))

        host_call()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            host_call()
        host_ms = (time.perf_counter() - t0) * 1e3 / args.steps
        # the schedule-free grid on the same device buffers: each cell's exact
        # FP32 nearest distance (no counterpart in the reference)
        dexact = torch.empty_like(dout)

        def exact_call():
            _lib.check(L.sdfx_generate_grid_sdf_exact_device(
                ctypes.c_void_p(dv.data_ptr()), len(v), ctypes.c_void_p(df.data_ptr()), df.numel(), 0,
                first_a.ctypes.data_as(F), size_a.ctypes.data_as(F),
                cnt_a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), 0, ctypes.c_void_p(dexact.data_ptr()), sp))

        exact_call()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            exact_call()
        torch.cuda.synchronize()
        exact_ms = (time.perf_counter() - t0) * 1e3 / args.steps
        forms = {"device_pointers": {"ms_per_call": ms, "samples_per_s": value},
                 "host_pointers": {"ms_per_call": host_ms, "samples_per_s": samples / host_ms * 1e3,
                                   "note": "sdfx_generate_grid_sdf: mesh upload and distance download included"},
                 "exact": {"ms_per_call": exact_ms, "samples_per_s": samples / exact_ms * 1e3,
                           "cells_differing_from_flood_fill": int((dexact != dout).sum().item()),
                           "note": "sdfx_generate_grid_sdf_exact_device: every cell the FP32 nearest distance, "
                                   "signs from the same grid raycasts"}}

    # Kernel timing: the same K steps again with HIP events around every launch
    # (on the launch stream), plus exact work counters.
    if args.no_kernel_timing:
        print(json.dumps({"value": value, "ms_per_step": ms, "steps": args.steps, "forms": forms}), flush=True)
        return
    _lib.profile_enable(True)
    timed(args.steps)
    prof = _lib.profile_read()
    counters = {k: _lib.profile_counter(k) for k in (
        "grid.seed_distance_evals", "grid.propagate_distance_evals", "lbvh.nearest_nodes", "lbvh.nearest_leaves",
        "lbvh.vote_nodes", "lbvh.vote_leaves", "lbvh.queries", "lbvh.nearest_wave_slots",
        "lbvh.vote_wave_slots", "lbvh.deferred_queries", "lbvh.deferred_nodes", "lbvh.deferred_leaves",
        "lbvh.nearest_drops", "lbvh.restarts")}
    _lib.profile_enable(False)
    kernels = sorted(((name, t, n) for name, (t, n) in prof.items()), key=lambda x: -x[1])
    dom_name = kernels[0][0]
    # The dominant kernel's average duration: the same K steps once more with an
    # event pair on that kernel's launches only and the work counters off.
    _lib.profile_enable(True)
    _lib.profile_select(dom_name, counters=False)
    timed(args.steps)
    dom_ms, dom_launches = _lib.profile_read()[dom_name]
    _lib.profile_enable(False)
    # Algorithmic bytes per SDF sample (implementation-independent): a grid cell
    # writes 4 B, a point query reads 12 B and writes 4 B; the mesh (12 B per
    # vertex and per index triple) is read once per call. One call runs
    # `launches_per_step` launches of the dominant kernel, which are charged
    # the call's bytes in equal parts.
    per_sample = 0.0 if args.workload == "mesh_check" else 4.0 if args.workload.startswith("grid") else 16.0
    work_samples, work_mesh = samples, mesh_bytes
    if args.workload == "serve":
        # each kernel is charged the requests it serves: the LBVH query kernels
        # the point requests (16 B per point plus their meshes), the grid
        # kernels the grid requests (4 B per cell plus their meshes)
        if "lbvh" in dom_name:
            per_sample, work_samples, work_mesh = 16.0, serve.point_samples, serve.point_mesh_bytes
        else:
            per_sample, work_samples, work_mesh = 4.0, serve.grid_samples, serve.grid_mesh_bytes
    step_bytes = per_sample * work_samples + work_mesh
    launches_per_step = dom_launches / args.steps
    avg_launch_s = dom_ms / dom_launches / 1e3
    achieved = step_bytes / launches_per_step / avg_launch_s / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s", "frac": achieved / 8000.0,
                "traffic": measured_traffic(args.workload, dom_name), "kernel": dom_name, "avg_launch_ms": dom_ms / dom_launches,
                "work_per_launch": f"{step_bytes / launches_per_step:.0f} B ({per_sample:.0f} B x {work_samples} "
                                   f"samples + {work_mesh} B of mesh, over {launches_per_step:g} launches per call)"}
    valu = None
    # FP32 VALU ceiling of this code: without FMA contraction (the reference's
    # rounding) and without packed FP32, one lane retires one FP32 operation
    # per instruction, a quarter of the 157.3 TFLOP/s peak (which counts an FMA
    # as two and packed pairs as two).
    valu_ceiling = 157.3 / 4
    if args.workload == "none" and dom_name.startswith("k_none_raycast"):
        # 116 FP32 operations per (query, triangle) pair: closest point and
        # distance (80) and the +X ray test (36), geo.rs as written
        pairs = float(len(q)) * len(f)
        tflops = 116.0 * pairs / launches_per_step / avg_launch_s / 1e12
        valu = {"achieved": tflops, "peak": 157.3, "unit": "TFLOP/s", "frac": tflops / 157.3,
                "frac_of_unfused_ceiling": tflops / valu_ceiling, "pairs_per_launch": pairs / launches_per_step}
    if args.workload.startswith("grid"):
        # diagnostic beside the headline: FP32 VALU work of the distance evaluations
        # (80 operations per point-triangle distance, geo.rs:26-30, 70-138 as written)
        evals = None
        if dom_name.startswith("k_frontier_propagate"):
            evals = counters["grid.propagate_distance_evals"]
        elif dom_name.startswith("k_seed_pairs"):
            evals = counters["grid.seed_distance_evals"]
        if evals:
            tflops = 80.0 * evals / dom_launches / avg_launch_s / 1e12
            valu = {"achieved": tflops, "peak": 157.3, "unit": "TFLOP/s", "frac": tflops / 157.3,
                    "frac_of_unfused_ceiling": tflops / valu_ceiling, "evals_per_launch": evals / dom_launches}

    cpu = None if args.no_cpu_baseline else cpu_baseline(args.workload, v, f, locals())
    if cpu is not None and "host" not in cpu:
        cpu["host"] = host_cpus()[1]

    checking = args.workload == "mesh_check"
    result = {
        "metric": "triangles checked/sec (1 GPU)" if checking else "SDF samples/sec (1 GPU)",
        "value": value,
        "unit": "triangles/s" if checking else "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded mesh generators)",
        "config": config,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels_ms_per_step": {n: round(t / args.steps, 4) for n, t, _ in kernels[:8]},
    }
    if forms:
        result["forms"] = forms
    if valu:
        result["valu_distance_evals"] = valu
    if counters["lbvh.queries"]:
        nq_total = counters["lbvh.queries"]
        result["traversal_per_query"] = {k.split(".", 1)[1]: counters[k] / nq_total for k in (
            "lbvh.nearest_nodes", "lbvh.nearest_leaves", "lbvh.vote_nodes", "lbvh.vote_leaves",
            "lbvh.nearest_drops", "lbvh.restarts") if k in counters}
        if counters.get("lbvh.deferred_queries"):
            # nearest searches past the visit budget, finished one query per wave
            nd = counters["lbvh.deferred_queries"]
            result["traversal_per_query"].update({
                "deferred_fraction": nd / nq_total,
                "deferred_nodes_per_deferred_query": counters["lbvh.deferred_nodes"] / nd,
                "deferred_leaves_per_deferred_query": counters["lbvh.deferred_leaves"] / nd})
        # SIMD efficiency: node visits / (64 x the slowest lane's visits, per wave)
        for k in ("nearest", "vote"):
            if counters[f"lbvh.{k}_wave_slots"]:
                result["traversal_per_query"][f"{k}_simd_efficiency"] = (
                    counters[f"lbvh.{k}_nodes"] / counters[f"lbvh.{k}_wave_slots"])
    if cpu and cpu.get("speedup") is None:
        cpu["speedup"] = value / cpu["value"] if cpu["value"] else None
    print(json.dumps(result), flush=True)


class ServeWorkload:
    """64 resident meshes and a stream of 256 point / grid requests on a server."""

    def __init__(self, torch, dev, L, sp, n_meshes=64, n_requests=256, workers=2, budget=32 << 30,
                 seed=0x5DF00005):
        import numpy as np
        from mesh_to_sdf_amd import _lib, synthetic as S
        self.torch, self.L, self.sp, self.np, self._lib = torch, L, sp, np, _lib
        self.workers, self.budget = workers, budget
        self.meshes = S.config4(n_meshes, seed=seed)
        h = ctypes.c_void_p()
        _lib.check(L.sdfx_server_create(budget, workers, ctypes.byref(h)))
        self.h = h
        self.dmesh = []
        self.mesh_bytes = 0
        descs = (_lib.ServerMesh * n_meshes)()
        for d, (mv, mf, *_) in zip(descs, self.meshes):
            dvi = torch.from_numpy(mv).to(dev)
            dfi = torch.from_numpy(mf.reshape(-1).astype(np.int32)).to(dev)
            self.dmesh.append((dvi, dfi))
            d.vertices, d.vertex_count, d.indices, d.index_count, d.topology = (
                dvi.data_ptr(), len(mv), dfi.data_ptr(), dfi.numel(), 0)
        torch.cuda.synchronize()
        # the pool is admitted with one batched LBVH build
        t0 = time.perf_counter()
        _lib.check(L.sdfx_server_add_meshes(h, descs, n_meshes))
        self.pool_build_ms = (time.perf_counter() - t0) * 1e3
        self.ids = [descs[i].id for i in range(n_meshes)]
        self.reqs = []
        self.point_samples = self.grid_samples = 0
        self.point_mesh_bytes = 0  # mesh bytes read by the point requests (12 B per vertex and per index triple)
        self.grid_mesh_bytes = 0
        for kind, mi, a in S.serve_requests(self.meshes, n_requests, seed=seed + 1):
            mv, mf = self.meshes[mi][0], self.meshes[mi][1]
            if kind == "points":
                n = len(a)
                out = torch.empty(n, dtype=torch.float32, device=dev)
                self.reqs.append(("points", mi, torch.from_numpy(a).to(dev), out, n))
                self.point_samples += n
                self.point_mesh_bytes += 12 * len(mv) + 12 * len(mf)
            else:
                first, size, count = a
                n = count[0] * count[1] * count[2]
                out = torch.empty(n, dtype=torch.float32, device=dev)
                grid = ((ctypes.c_float * 3)(*first), (ctypes.c_float * 3)(*size), (ctypes.c_uint64 * 3)(*count))
                self.reqs.append(("grid", mi, grid, out, n))
                self.grid_samples += n
                self.grid_mesh_bytes += 12 * len(mv) + 12 * len(mf)
        torch.cuda.synchronize()
        self.samples = self.point_samples + self.grid_samples

    def submit(self, r):
        kind, mi, a, out, n = r
        tk = ctypes.c_uint64(0)
        if kind == "points":
            st = self.L.sdfx_server_submit_points(self.h, self.ids[mi], a.data_ptr(), n, 1, 0, out.data_ptr(),
                                                  ctypes.byref(tk))
        else:
            st = self.L.sdfx_server_submit_grid(self.h, self.ids[mi], a[0], a[1], a[2], 0, out.data_ptr(),
                                                ctypes.byref(tk))
        self._lib.check(st)
        return tk.value

    def wait(self, tk):
        ms = ctypes.c_double(0.0)
        self._lib.check(self.L.sdfx_server_wait(self.h, tk, ctypes.byref(ms)))
        return ms.value

    def burst(self):
        tks = [self.submit(r) for r in self.reqs]
        return [self.wait(t) for t in tks]

    def poisson(self, lam, seed):
        g = self.np.random.Generator(self.np.random.PCG64(seed))
        gaps = g.exponential(1.0 / lam, size=len(self.reqs))
        tks = []
        t0 = time.perf_counter()
        at = t0
        for r, gap in zip(self.reqs, gaps):
            at += gap
            while time.perf_counter() < at:
                pass
            tks.append(self.submit(r))
        lat = [self.wait(t) for t in tks]
        return time.perf_counter() - t0, lat

    def stats(self):
        s = self._lib.ServerStats()
        self._lib.check(self.L.sdfx_server_get_stats(self.h, ctypes.byref(s)))
        return {n: int(getattr(s, n)) for n, _ in self._lib.ServerStats._fields_}

    def plain(self):
        """The plain drop-in, one request at a time (each point call builds its LBVH)."""
        F = ctypes.POINTER(ctypes.c_float)
        U = ctypes.POINTER(ctypes.c_uint64)
        lat = []
        t_all = time.perf_counter()
        for kind, mi, a, out, n in self.reqs:
            dvi, dfi = self.dmesh[mi]
            t0 = time.perf_counter()
            if kind == "points":
                cnt = ctypes.c_uint64(0)
                st = self.L.sdfx_generate_sdf_device(dvi.data_ptr(), dvi.shape[0], dfi.data_ptr(), dfi.numel(), 0,
                                                     a.data_ptr(), n, 1, 0, out.data_ptr(), ctypes.byref(cnt), self.sp)
            else:
                st = self.L.sdfx_generate_grid_sdf_device(dvi.data_ptr(), dvi.shape[0], dfi.data_ptr(), dfi.numel(),
                                                          0, ctypes.cast(a[0], F), ctypes.cast(a[1], F),
                                                          ctypes.cast(a[2], U), 0, out.data_ptr(), self.sp)
            self._lib.check(st)
            lat.append((time.perf_counter() - t0) * 1e3)
        return time.perf_counter() - t_all, lat

    def forms(self, value, ms, step_ms):
        np = self.np
        pct = lambda a, p: float(np.percentile(np.asarray(a), p))
        out = {"served_burst": {"samples_per_s": value, "ms_per_stream": ms,
                                "note": "all requests submitted at once; saturated throughput"}}
        sweep = []
        for lam in SERVE_LAMBDAS:
            b0 = self.stats()
            span, lat = self.poisson(lam, 1000 + lam)
            b1 = self.stats()
            nb = b1["batches"] - b0["batches"]
            sweep.append({"lambda_req_per_s": lam, "samples_per_s": self.samples / span, "seconds": span,
                          "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99), "batches": nb,
                          "requests_per_batch": len(self.reqs) / max(1, nb)})
        out["served_poisson"] = sweep
        self.plain()  # warm the plain path's allocations
        span, lat = self.plain()
        out["plain_sequential"] = {"samples_per_s": self.samples / span, "seconds": span, "p50_ms": pct(lat, 50),
                                   "p99_ms": pct(lat, 99),
                                   "note": "sdfx_generate_sdf_device / sdfx_generate_grid_sdf_device per request"}
        out["served_over_plain"] = value / (self.samples / span)
        out["server"] = self.stats()
        out["pool_admission_ms"] = self.pool_build_ms
        out["pool_admission"] = f"{len(self.meshes)} meshes, one sdfx_server_add_meshes call (batched LBVH build)"
        return out


def host_cpus():
    """(threads the CPU baseline uses, host details). The threads are the CPUs
    this process may run on, limited by OMP_NUM_THREADS when the machine sets
    it (a GPU machine sets it to the CPU share it gives one GPU)."""
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = os.environ.get("OMP_NUM_THREADS", "")
    threads = max(1, min(avail, int(share))) if share.isdigit() and int(share) > 0 else avail
    model = governor = None
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), None)
    except OSError:
        pass
    try:
        with open("/sys/devices/system/cpu/cpu0/cpufreq/scaling_governor") as fh:
            governor = fh.read().strip()
    except OSError:
        pass
    return threads, {"nproc": os.cpu_count(), "affinity_cpus": avail, "omp_num_threads": share or None,
                     "threads_used": threads, "cpu_model": model, "governor": governor}


def serve_accuracy(sv, O, S, np, threads, budget_pairs=3e9):
    """Accuracy of the served answers on a bounded sample, against the oracle
    and an FP64 brute-force distance (Ericson's closest point in double over
    every triangle). Points: the first point requests' answers (a prefix of
    each, so the FP64 work stays within budget_pairs point-triangle pairs)
    against the oracle's Bvh(Raycast). Grid: the grid request with the fewest
    cell x triangle pairs against the reference's single-thread heap schedule
    restated by the oracle. Samples whose FP64 distance is within 1e-5 x bbox
    diagonal of the surface are left out of the sign comparison and counted."""
    pts = [r for r in sv.reqs if r[0] == "points"][:4]
    per_req = budget_pairs / 2 / max(1, len(pts))
    gpu_e, cpu_e, near, sign_bad, n_tot, dmax = [], [], 0, 0, 0, 0.0
    for kind, mi, dq, out, n in pts:
        mv, mf = sv.meshes[mi][0], sv.meshes[mi][1]
        k = int(max(1, min(n, per_req / len(mf))))
        q = dq[:k].cpu().numpy()
        gpu = out[:k].cpu().numpy()
        s, ref = O.generate_sdf(mv, mf, O.TOPO_LIST, q, O.ACCEL_BVH, O.SIGN_RAYCAST, threads)
        d64 = O.fp64_distance(mv, mf, q, threads)
        diag = S.bbox_diagonal(mv)
        far = d64 > 1e-5 * diag
        near += int(np.sum(~far))
        sign_bad += int(np.sum(np.signbit(gpu[far]) != np.signbit(ref[far])))
        dmax = max(dmax, float(np.max(np.abs(gpu - ref)) / diag))
        gpu_e.append(np.abs(np.abs(gpu).astype(np.float64) - d64) / diag)
        cpu_e.append(np.abs(np.abs(ref).astype(np.float64) - d64) / diag)
        n_tot += k
    ge, ce = np.concatenate(gpu_e), np.concatenate(cpu_e)
    points = {"samples": n_tot, "requests": len(pts), "reference": "oracle Bvh(Raycast) (generic/bvh.rs)",
              "sign_mismatches_excluding_near_surface": sign_bad, "near_surface_excluded": near,
              "max_abs_diff_over_diag": dmax,
              "fp64_error_over_diag": {"gpu_max": float(ge.max()), "gpu_p99": float(np.percentile(ge, 99)),
                                       "cpu_max": float(ce.max()), "cpu_p99": float(np.percentile(ce, 99))}}
    grids = [r for r in sv.reqs if r[0] == "grid"]
    kind, mi, a, out, n = min(grids, key=lambda r: r[4] * len(sv.meshes[r[1]][1]))
    mv, mf = sv.meshes[mi][0], sv.meshes[mi][1]
    first, size, count = np.array(a[0][:], np.float32), np.array(a[1][:], np.float32), tuple(int(c) for c in a[2])
    s, ref, _ = O.generate_grid_sdf(mv, mf, O.TOPO_LIST, first, size, count, O.SIGN_RAYCAST, 1, O.GRID_SEQUENTIAL)
    gpu = out.cpu().numpy()
    cells = O.grid_centres(first, size, count)
    d64 = O.fp64_distance(mv, mf, cells, threads)
    diag = S.bbox_diagonal(mv)
    far = d64 > 1e-5 * diag
    ge = np.abs(np.abs(gpu).astype(np.float64) - d64) / diag
    ce = np.abs(np.abs(ref).astype(np.float64) - d64) / diag
    dd = np.abs(gpu - ref) / diag
    grid = {"cells": int(n), "triangles": int(len(mf)),
            "reference": "oracle restatement of generate/grid.rs, 1-thread heap schedule",
            "sign_mismatches_excluding_near_surface": int(np.sum(np.signbit(gpu[far]) != np.signbit(ref[far]))),
            "near_surface_excluded": int(np.sum(~far)),
            "cells_beyond_1e-5_diag_vs_reference": int(np.sum(dd > 1e-5)), "max_abs_diff_over_diag": float(dd.max()),
            "fp64_error_over_diag": {"gpu_max": float(ge.max()), "gpu_p99": float(np.percentile(ge, 99)),
                                     "cpu_max": float(ce.max()), "cpu_p99": float(np.percentile(ce, 99)),
                                     "gpu_cells_beyond_1e-5": int(np.sum(ge > 1e-5)),
                                     "cpu_cells_beyond_1e-5": int(np.sum(ce > 1e-5))}}
    return {"points": points, "grid": grid}


def cpu_baseline(workload, v, f, env):
    """The oracle's restatement of the reference's CPU algorithm on this host."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    from mesh_to_sdf_amd import synthetic as S

    threads, host = host_cpus()
    budget_s, times = 30.0, []
    if workload == "grid":
        first, size, count = env["first"], env["size"], env["count"]
        samples = count[0] * count[1] * count[2]
        t_start = time.perf_counter()
        while len(times) < 3 and time.perf_counter() - t_start < budget_s:
            t0 = time.perf_counter()
            s, out, _ = O.generate_grid_sdf(v, f, O.TOPO_LIST, first, size, count, O.SIGN_RAYCAST, threads,
                                            O.GRID_THREADED)
            times.append(time.perf_counter() - t0)
        t = sorted(times)[len(times) // 2]
        gpu = env["dout"].cpu().numpy()
        diag = S.bbox_diagonal(v)
        diff = np.abs(gpu - out)
        accuracy = {"sign_mismatches_vs_reference_schedule": int(np.sum(np.signbit(gpu) != np.signbit(out))),
                    "cells_differing": int(np.sum(diff > 0)), "max_abs_diff_over_diag": float(diff.max() / diag)}
        return {"value": samples / t, "unit": "samples/s", "cores": threads, "kind": "port",
                "sample": f"full workload ({len(f)} tris, {samples} cells); the oracle's restatement of "
                          f"generate/grid.rs with the heap split over {threads} threads; median of {len(times)}",
                "seconds": t, "gpu_vs_this_run": accuracy, "speedup": None}
    if workload == "grid_batch":
        # a bounded sample: the requests in order until about 20 s of CPU time
        cells = 0
        t_start = time.perf_counter()
        used = 0
        for mv, mf, mfirst, msize, mcount in env["meshes"]:
            if time.perf_counter() - t_start > 20.0:
                break
            s, _, _ = O.generate_grid_sdf(mv, mf, O.TOPO_LIST, mfirst, msize, mcount, O.SIGN_RAYCAST, threads,
                                          O.GRID_THREADED)
            cells += mcount[0] * mcount[1] * mcount[2]
            used += 1
        t = time.perf_counter() - t_start
        return {"value": cells / t, "unit": "samples/s", "cores": threads, "kind": "port",
                "sample": f"the first {used} of the {len(env['meshes'])} requests ({cells} cells), one call each; "
                          f"the oracle's restatement of generate/grid.rs with the heap split over {threads} threads",
                "seconds": t, "speedup": None}
    if workload == "mesh_check":
        # the oracle restates the reference's O(n^2) scans as written: a smaller
        # mesh of the same family keeps it bounded
        sv, sf = S.icosphere(4)
        t0 = time.perf_counter()
        k, _ = O.check_mesh_triangle_list_at(sv, sf)
        t = time.perf_counter() - t0
        assert k == 0
        return {"value": len(sf) / t, "unit": "triangles/s", "cores": 1, "kind": "port",
                "sample": f"icosphere L4 ({len(sf)} triangles): the reference's scans are O(n^2) "
                          "(generate/mesh_check.rs:34-77), so the full mesh is out of reach; the rate falls "
                          "linearly with the mesh size", "seconds": t, "speedup": None}
    if workload == "serve":
        # a bounded sample: the request stream in order until about 20 s of CPU
        # time, each request through the oracle's restatement of its reference path
        sv = env["serve"]
        done = used = 0
        t_start = time.perf_counter()
        for kind, mi, a, _, n in sv.reqs:
            if time.perf_counter() - t_start > 20.0:
                break
            mv, mf = sv.meshes[mi][0], sv.meshes[mi][1]
            if kind == "points":
                O.generate_sdf(mv, mf, O.TOPO_LIST, a.cpu().numpy(), O.ACCEL_BVH, O.SIGN_RAYCAST, threads)
            else:
                O.generate_grid_sdf(mv, mf, O.TOPO_LIST, np.array(a[0][:], np.float32), np.array(a[1][:], np.float32),
                                    tuple(int(c) for c in a[2]), O.SIGN_RAYCAST, threads, O.GRID_THREADED)
            done += n
            used += 1
        t = time.perf_counter() - t_start
        return {"value": done / t, "unit": "samples/s", "cores": threads, "kind": "port",
                "sample": f"the first {used} of the {len(sv.reqs)} requests ({done} samples) through the oracle's "
                          f"restatement of generic/bvh.rs (points) and generate/grid.rs (grids), {threads} threads",
                "seconds": t, "speedup": None, "host": host,
                "accuracy": serve_accuracy(sv, O, S, np, threads)}
    if workload == "rtree_bvh":
        q = env["q"]
        t_start = time.perf_counter()
        while len(times) < 5 and time.perf_counter() - t_start < budget_s:
            t0 = time.perf_counter()
            O.generate_sdf(v, f, O.TOPO_LIST, q, O.ACCEL_RTREE_BVH, O.SIGN_RAYCAST, 1)
            times.append(time.perf_counter() - t0)
        t = sorted(times)[len(times) // 2]
        return {"value": len(q) / t, "unit": "samples/s", "cores": 1, "kind": "port",
                "sample": "full workload; single-threaded like generic/rtree_bvh.rs:150-152", "seconds": t,
                "speedup": None}
    if workload == "none":
        q = env["q"][:65_536]
        t0 = time.perf_counter()
        O.generate_sdf(v, f, O.TOPO_LIST, q, O.ACCEL_NONE, O.SIGN_RAYCAST, threads)
        t = time.perf_counter() - t0
        return {"value": len(q) / t, "unit": "samples/s", "cores": threads, "kind": "port",
                "sample": f"first 65,536 of the 1,048,576 queries, {threads} threads (default.rs par_iter)",
                "seconds": t, "speedup": None}
    accel = O.ACCEL_RTREE if workload == "rtree" else O.ACCEL_BVH
    q = env["q"][:200_000]
    t0 = time.perf_counter()
    O.generate_sdf(v, f, O.TOPO_LIST, q, accel, O.SIGN_NORMAL if workload == "rtree" else O.SIGN_RAYCAST, threads)
    t = time.perf_counter() - t0
    return {"value": len(q) / t, "unit": "samples/s", "cores": threads, "kind": "port",
            "sample": f"first 200,000 of the 4,194,304 queries, {threads} threads (rayon par_iter in the reference)",
            "seconds": t, "speedup": None}


if __name__ == "__main__":
    main()
