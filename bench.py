#!/usr/bin/env python3
"""Benchmark: seconds per hemisphere of CAT_SurfWarp on one MI355X.

Workload (fixed): a 61,442-vertex hemisphere pair built as the convex hull of
a seeded jittered Fibonacci sphere (jitter 0.2 of spacing), degree-16
spherical-harmonic features with N(0, 0.05) noise, target rotated (<= 30 deg)
and smoothly warped, registered with steps=2 runs=2 avg, for both the LH and
the x-mirrored RH copy.  One timed step = LH + RH.  Inputs are resident in
memory (no file I/O) when the timed region starts.

Prints one JSON line (rank 0).  Besides the wall-clock value it reports the
roofline of the dominant hot-path kernel (the timed kernel with the largest
total time): the library brackets each of its launches inside the timed
region with HIP events on the stream it runs on and accumulates the launch's
algorithmic bytes.  A CPU baseline from the project's C restatement of the
DARTEL step (oracle/) is timed on a bounded sample on the host.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "cat-surface-gpu_amd"))
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def make_inputs(seed: int):
    from cat_surface_gpu import synthetic
    from cat_surface_gpu.gpu_pipeline import GiftiMesh

    pairs = []
    for mirror in (False, True):
        p = synthetic.make_pair("fibonacci", 61442, lmax=16, noise=0.05, max_rotation_deg=30.0,
                                warp_deg=3.0, seed=seed, mirror=mirror)
        pairs.append((GiftiMesh(p.src_surface, p.src_faces), GiftiMesh(p.src_sphere, p.src_faces),
                      GiftiMesh(p.tgt_surface, p.tgt_faces), GiftiMesh(p.tgt_sphere, p.tgt_faces)))
    return pairs


def run_step(pairs, dtype: str, mode: str, device, cuda_graph: bool):
    """One step: LH and RH.  "plain" registers them one after the other (the
    reference's pipeline, one job per call); "batched" schedules both jobs
    together (run_cat_surface_gpu_pipelines: shared feature solve and batched
    DARTEL rounds, identical results); "batched+adjoint" also replaces the
    Nelder-Mead rotation refinement by gradient (Levenberg-Marquardt) steps on
    the 8 best seeds."""

    from cat_surface_gpu import (PipelineJob, run_cat_surface_gpu_pipeline,
                                 run_cat_surface_gpu_pipelines)

    kw = dict(device=device, dartel_dtype=dtype, steps=2, runs=2, avg=True, write_output=False,
              cuda_graph=cuda_graph)
    if mode == "batched+adjoint":
        kw.update(rotation_method="adjoint")
    if mode in ("batched", "batched+adjoint"):
        return run_cat_surface_gpu_pipelines([PipelineJob(meshes=m) for m in pairs], **kw)
    return [run_cat_surface_gpu_pipeline(None, None, None, None, "unused.gii", meshes=m, **kw)
            for m in pairs]


def kernel_rooflines() -> list[dict]:
    """Totals of the library's HIP-event kernel timers since they were enabled:
    for each timed kernel, its launches, average launch duration and the
    algorithmic bytes of those launches (what the reference's operator sequence
    reads and writes once for the same work) over the measured time."""

    import ctypes

    from cat_surface_gpu import _lib

    lib = _lib.lib()
    out = []
    k = 0
    while True:
        name = lib.csg_kernel_timing_name(k)
        if not name:
            break
        ms, nbytes, launches = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
        _lib.check(lib.csg_kernel_timing_read(k, ctypes.byref(ms), ctypes.byref(nbytes),
                                              ctypes.byref(launches)), "kernel timing")
        if launches.value:
            achieved = nbytes.value / (ms.value / 1e3) / 1e9
            out.append({"kernel": name.decode(), "bound": "hbm", "achieved": round(achieved, 2),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                        "launches": launches.value, "total_ms": round(ms.value, 3),
                        "avg_launch_us": round(ms.value * 1e3 / launches.value, 3),
                        "algorithmic_bytes_per_launch": round(nbytes.value / launches.value)})
        k += 1
    return sorted(out, key=lambda r: -r["total_ms"])


def cpu_baseline() -> dict:
    """Project C restatement (oracle) of dartel_step at 512x256 float64, one step
    per k = 0..5, single thread; scaled to the 108 dartel steps of one hemisphere
    (3 solves x 2 steps x 6 loops x 3 its).  DARTEL only: a lower bound of a full
    CPU hemisphere."""

    import numpy as np

    from oracle import oracle as O

    nx, ny = 512, 256
    x = (np.arange(nx) + 0.5)[None, :] / nx
    y = (np.arange(ny) + 0.5)[:, None] / ny
    src = 0.5 + 0.3 * np.sin(2 * np.pi * 3 * x) * np.cos(2 * np.pi * y)
    tgt = np.roll(src, 3, axis=1)
    v = np.zeros((2, ny, nx))
    per_k = []
    params = O.default_dartel_parameters(6)
    for k in range(6):
        t = time.perf_counter()
        v, _ = O.dartel_step(src, tgt, v, k, params[k])
        per_k.append(time.perf_counter() - t)
    seconds = 3 * 2 * 3 * sum(per_k)
    return {"value": round(seconds, 3), "unit": "s/hemisphere (DARTEL stage only)", "cores": 1,
            "kind": "port",
            "sample": "oracle dartel_step 512x256 fp64, one step for each k=0..5 "
                      f"({sum(per_k):.2f} s), x18 for 3 solves x 2 steps x 3 its"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--dtype", choices=("float32", "float64"), default="float32")
    ap.add_argument("--mode", choices=("plain", "batched", "batched+adjoint"),
                    default="batched+adjoint")
    ap.add_argument("--seed", type=int, default=20261205)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--compare-plain", action=argparse.BooleanOptionalAction, default=True,
                    help="after the timed mode, time the plain path on the same inputs and "
                         "report it with the largest vertex difference between the two")
    ap.add_argument("--cuda-graph", action=argparse.BooleanOptionalAction, default=True,
                    help="replay the DARTEL solves from captured graphs (the kernels inside a "
                         "replay are not seen by the kernel timers)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: every registration here fits one GPU and this "
              "build runs no multi-GPU path (--gpus must be 1)", file=sys.stderr)
        return 2
    import torch

    device = torch.device("cuda", int(os.environ.get("LOCAL_RANK", "0")))
    torch.cuda.set_device(device)
    t = time.perf_counter()
    pairs = make_inputs(args.seed)
    log(f"inputs ready in {time.perf_counter() - t:.1f} s")
    for i in range(args.warmup):
        run_step(pairs, args.dtype, args.mode, device, args.cuda_graph)
        torch.cuda.synchronize()
        log(f"warmup {i} done")
    from cat_surface_gpu import _lib

    kernel_rooflines()  # clear
    torch.cuda.synchronize()
    _lib.lib().csg_kernel_timing(1)
    t0 = time.perf_counter()
    stage_sums: dict[str, float] = {}
    for i in range(args.steps):
        results = run_step(pairs, args.dtype, args.mode, device, args.cuda_graph)
        # batched jobs share their stage timings: count them once
        for r in (results[:1] if args.mode != "plain" else results):
            for key, val in r.timings.items():
                stage_sums[key] = stage_sums.get(key, 0.0) + val
        log(f"step {i} done")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    _lib.lib().csg_kernel_timing(0)
    ms_per_step = elapsed / args.steps * 1e3
    hemispheres = 2
    value = elapsed / (args.steps * hemispheres)
    kernels = kernel_rooflines()
    roof = kernels[0] if kernels else None
    plain = None
    if args.compare_plain and args.mode != "plain":
        for i in range(args.warmup):
            run_step(pairs, args.dtype, "plain", device, args.cuda_graph)
        torch.cuda.synchronize()
        tp = time.perf_counter()
        for i in range(args.steps):
            plain_results = run_step(pairs, args.dtype, "plain", device, args.cuda_graph)
        torch.cuda.synchronize()
        plain_value = (time.perf_counter() - tp) / (args.steps * hemispheres)
        delta = max(float(abs(a.vertices.astype("float64") - b.vertices).max())
                    for a, b in zip(results, plain_results))
        plain = {"value": round(plain_value, 4), "unit": "s/hemisphere",
                 "speedup": round(plain_value / value, 3),
                 "max_vertex_delta": delta}
        log("plain comparison done")
    cpu = None if args.no_cpu_baseline else cpu_baseline()
    stages = {k: round(v / (args.steps * hemispheres), 4) for k, v in sorted(stage_sums.items())}
    line = {
        "metric": "s/hemisphere CAT_SurfWarp 61,442 verts (steps=2 runs=2 avg)",
        "value": round(value, 4),
        "unit": "s/hemisphere",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.dtype == "float32" else "f64",
        "data": "synthetic (seeded Fibonacci-hull hemispheres, SH features)",
        "config": {"workload": "61,442-vertex hemisphere pair, LH+RH per step, steps=2 runs=2 "
                               "avg, 512x256 map", "mode": args.mode,
                   "parallelism": "single GPU", "cuda_graph": args.cuda_graph},
        "plain": plain,
        "roofline": roof,
        "roofline_kernels": kernels,
        "cpu_baseline": cpu,
        "stages_s_per_hemisphere": stages,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
