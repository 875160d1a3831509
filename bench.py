"""SDF samples/s of the MI355X path on one GPU.

Default workload (the metric's grid configuration): generate_grid_sdf with
SignMethod::Raycast on a jittered torus (R=1, r=0.35, 256x128 segments, 65,536
triangles) over a 128^3 grid built from its AABB expanded by 10%; one step =
one full generate_grid_sdf call (topology decode, seeding, propagation to the
fixpoint, raycast sign, output write) on vertex/index buffers already resident
in HBM. One SDF sample = one output cell.

    python bench.py [--steps K] [--warmup W] [--workload grid|rtree_bvh|bvh]

Prints one JSON line. The CPU baseline is the oracle's restatement of the
reference's threaded algorithm, timed on this host's cores.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOADS = {
    "grid": "generate_grid_sdf Raycast, jittered torus 65,536 tris, 128^3 cells",
    "rtree_bvh": "generate_sdf RtreeBvh, icosphere L3 1,280 tris, 10,000 uniform queries",
    "bvh": "generate_sdf Bvh(Raycast), noise sphere L7 327,680 tris, 4,194,304 queries",
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="grid", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    return ap.parse_args()


def main():
    args = parse()
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
        config = {"workload": WORKLOADS["grid"], "triangles": int(len(f)), "cells": int(ncells),
                  "sign": "Raycast"}
    else:
        if args.workload == "rtree_bvh":
            v, f, q = S.config1()
            method = (3, 0)
        else:
            v, f, q = S.config3()
            method = (1, 0)
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
        config = {"workload": WORKLOADS[args.workload], "triangles": int(len(f)), "queries": int(len(q))}

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms = elapsed * 1000.0 / args.steps
    value = samples * args.steps / elapsed

    result = {
        "metric": "SDF samples/sec (1 GPU)",
        "value": value,
        "unit": "samples/s",
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
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
