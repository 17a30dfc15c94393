#!/usr/bin/env python3
"""Benchmark: seconds per hemisphere of CAT_SurfWarp on one MI355X.

Workload (fixed): a 61,442-vertex hemisphere pair built as the convex hull of
a seeded jittered Fibonacci sphere (jitter 0.2 of spacing), degree-16
spherical-harmonic features with N(0, 0.05) noise, target rotated (<= 30 deg)
and smoothly warped, registered with steps=2 runs=2 avg, for both the LH and
the x-mirrored RH copy.  One timed step = LH + RH.  Inputs are resident in
memory (no file I/O) when the timed region starts.  The default mode runs the
batched+adjoint schedule as one captured HIP graph (mode "graph"): the first
call plans it eagerly and captures it, each later call copies the step's
meshes into the graph's input buffers, replays every kernel of the schedule
and checks its device-side verdicts (any failed check reruns the step eagerly).

Prints one JSON line (rank 0).  Besides the wall-clock value it reports:
- the roofline of the dominant hot-path kernel (the timed kernel with the
  largest total time) in one un-captured step of the timed mode run right
  after the timed region: the library brackets each launch with HIP events on
  the stream it runs on and accumulates the launch's algorithmic bytes (graph
  replays, which the timed region uses, hide their kernels from host events);
- the same mode with GIFTI input and output around each step (the files are
  read into memory, registered, and the output spheres written, all timed);
- the plain method on the same inputs;
- a CPU baseline: one full hemisphere on the project's CPU restatement
  (oracle/pipeline.py), wall and CPU time, cores stated.
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

# HBM bytes per launch of kernels whose traffic counters were collected at the
# bench's shape: two rocprofv3 --pmc runs of this bench (FETCH_SIZE, WRITE_SIZE),
# each scaled by what the counter reads for a 64 MiB copy, averaged over the run's
# dispatches.  Source: the named profile.
_TRAFFIC_PROFILE = "profiles/hbm_traffic_fibonacci61442_f32.md"
MEASURED_TRAFFIC = {
    "k_relax_tile": (13.21e6, _TRAFFIC_PROFILE),
    "k_fmg_coarse": (0.84e6, _TRAFFIC_PROFILE),
    "k_restrict_residual": (11.11e6, _TRAFFIC_PROFILE),
    "k_squaring": (69.26e6, _TRAFFIC_PROFILE),
    "k_objective": (66.15e6, _TRAFFIC_PROFILE),
    "k_compose_self": (32.11e6, _TRAFFIC_PROFILE),
    "k_combine": (53.59e6, _TRAFFIC_PROFILE),
    "k_smooth_geometry_patch": (14.39e6, _TRAFFIC_PROFILE),
    "k_sor_patched": (35.84e6, _TRAFFIC_PROFILE),
}


# What bounds a kernel whose algorithmic bytes (the reference operators' reads
# and writes) say little about its HBM use.
KERNEL_NOTES = {
    "k_sor_patched": "1000 colour-ordered sweeps; each sweep's 6 colour phases run on LDS "
                     "patches and the cores are exchanged through the XCD's L2 once per "
                     "sweep: bound by the latency of the 1000 exchange rounds (~6.6 us each), "
                     "not by HBM (hbm_frac)",
    "k_fmg_coarse": "all coarse multigrid levels of one cycle in one workgroup per job: "
                    "bound by ~100 dependent LDS phases, not by HBM",
}


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


def resident_inputs(pairs, device):
    """The same meshes as device tensors (the timed region starts with its inputs
    in HBM; the host-input leg reports the rate with the host copies)."""

    import torch

    from cat_surface_gpu.gpu_pipeline import GiftiMesh

    out = [tuple(GiftiMesh(torch.as_tensor(m.vertices, device=device),
                           torch.as_tensor(m.faces, device=device)) for m in pair)
           for pair in pairs]
    torch.cuda.synchronize(device)
    return out


def host_inputs_leg(pairs, args, device) -> dict:
    """The timed mode fed host arrays each step: the rate including the copies of
    the four input meshes per hemisphere into device memory and the faces'
    content hashing of the device cache."""

    import torch

    for _ in range(max(1, args.warmup)):
        run_step(pairs, args.dtype, args.mode, device, args.cuda_graph)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        run_step(pairs, args.dtype, args.mode, device, args.cuda_graph)
    torch.cuda.synchronize()
    value = (time.perf_counter() - t) / (args.steps * len(pairs))
    log("host-input leg done")
    return {"value": round(value, 4), "unit": "s/hemisphere",
            "definition": "the timed mode with the four input meshes of each hemisphere "
                          "handed over as host arrays every step (host-to-device copies "
                          "inside the timed region)"}


MODE_DEFINITIONS = {
    "plain": "one registration per call, 46 seeds + Nelder-Mead, per-operator multigrid, "
             "eager solves",
    "plain-fused": "plain with the fused multigrid schedule and graph-replayed solves",
    "batched": "both hemispheres per stage, batched solves replayed from per-shape graphs",
    "batched+adjoint": "batched, tuned relaxation tiles, top-8 seeds refined by 20 gradient "
                       "(Levenberg-Marquardt) steps; per-shape solve graphs",
    "graph": "batched+adjoint captured as one HIP graph (every stage), replayed per step on "
             "the step's meshes with device-side checks",
}

PLAIN_DEFINITION = ("one registration per call (the reference's pipeline), 46 seeds + "
                    "Nelder-Mead, one launch per reference multigrid operator (per-operator "
                    "FMG schedule), solves launched eagerly (no graph capture)")


def run_step(pairs, dtype: str, mode: str, device, cuda_graph: bool):
    """One step: LH and RH.

    "plain" is the plain method: the reference's pipeline one registration at a
    time, Nelder-Mead rotation refinement, and the DARTEL solves as one launch
    per reference operator (per-operator multigrid schedule), launched eagerly.
    "plain-fused" is the same with the fused multigrid schedule and graph replay.
    "batched" schedules both jobs together (run_cat_surface_gpu_pipelines:
    shared feature solve and batched DARTEL rounds on the fused schedule, graph
    replay, identical results); "batched+adjoint" also tunes the relaxation
    tiles and replaces Nelder-Mead by gradient (Levenberg-Marquardt) steps on
    the 8 best seeds; "graph" runs the batched+adjoint schedule as one captured
    graph (run_graph_schedule: every stage, device-side loop counts and checks,
    replayed each step)."""

    from cat_surface_gpu import (PipelineJob, run_cat_surface_gpu_pipeline,
                                 run_cat_surface_gpu_pipelines, set_fmg_schedule)

    kw = dict(device=device, dartel_dtype=dtype, steps=2, runs=2, avg=True, write_output=False,
              cuda_graph=cuda_graph)
    if mode in ("batched+adjoint", "graph"):
        kw.update(rotation_method="adjoint", autotune=True)
    if mode == "graph":
        # the whole batched+adjoint schedule as one graph: the first call plans
        # and captures it, later calls replay it
        kw.update(graph_schedule=cuda_graph)
    if mode in ("batched", "batched+adjoint", "graph"):
        return run_cat_surface_gpu_pipelines([PipelineJob(meshes=m) for m in pairs], **kw)
    if mode == "plain":
        kw.update(cuda_graph=False)
        previous = set_fmg_schedule(False)
        try:
            return [run_cat_surface_gpu_pipeline(None, None, None, None, "unused.gii",
                                                 meshes=m, **kw) for m in pairs]
        finally:
            set_fmg_schedule(previous)
    return [run_cat_surface_gpu_pipeline(None, None, None, None, "unused.gii", meshes=m, **kw)
            for m in pairs]


def _installed_tiles() -> list[dict]:
    from cat_surface_gpu import autotune

    return autotune.installed()


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
            traffic = MEASURED_TRAFFIC.get(name.decode())
            avg_us = ms.value * 1e3 / launches.value
            entry = {"kernel": name.decode(), "bound": "hbm", "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": traffic[0] if traffic else None,
                     "traffic_source": traffic[1] if traffic else None,
                     "launches": launches.value, "total_ms": round(ms.value, 3),
                     "avg_launch_us": round(avg_us, 3),
                     "algorithmic_bytes_per_launch": round(nbytes.value / launches.value)}
            if traffic:
                # the measured HBM bytes over the same launch time: what the kernel
                # actually draws from HBM (below the algorithmic bytes where its
                # working set stays in LDS and L2)
                hbm = traffic[0] / (avg_us * 1e-6) / 1e9
                entry.update(hbm_gbs=round(hbm, 1), hbm_frac=round(hbm / HBM_PEAK_GBS, 5))
            note = KERNEL_NOTES.get(name.decode())
            if note:
                entry["note"] = note
            out.append(entry)
        k += 1
    return sorted(out, key=lambda r: -r["total_ms"])


def cpu_baseline(pair) -> dict:
    """The project's CPU restatement (oracle/pipeline.py) of one full hemisphere:
    the same left-hemisphere inputs as the timed GPU step, all stages (stencils,
    rotation features, 46 seeds + Nelder-Mead, three float64 DARTEL solves with
    their curvature sheets, warps, averaging), on the host's cores (OpenMP over
    the oracle's elementwise loops; the float64 sums stay sequential).  Wall time,
    and user + system CPU time from getrusage."""

    from oracle import oracle as O
    from oracle import pipeline as P

    src_surface, src_sphere, tgt_surface, tgt_sphere = pair
    res, wall, cpu = P.timed_surfwarp(src_surface.vertices, src_sphere.vertices,
                                      src_sphere.faces, tgt_surface.vertices,
                                      tgt_sphere.vertices, tgt_sphere.faces, steps=2, runs=2,
                                      avg=True)
    threads = int(O.lib().oracle_max_threads())
    return {"value": round(wall, 3), "unit": "s/hemisphere", "cores": threads, "kind": "port",
            "cpu_seconds": round(cpu, 2),
            "stages_s": {k: round(v, 3) for k, v in res.stage_seconds.items()},
            "sample": "one full CAT_SurfWarp hemisphere (LH of the timed pair; steps=2 runs=2 "
                      "avg; float64) on the project CPU restatement, inputs in memory"}


def _pair_meshes(p):
    from cat_surface_gpu.gpu_pipeline import GiftiMesh

    return (GiftiMesh(p.src_surface, p.src_faces), GiftiMesh(p.src_sphere, p.src_faces),
            GiftiMesh(p.tgt_surface, p.tgt_faces), GiftiMesh(p.tgt_sphere, p.tgt_faces))


def uneven_report(seed: int, device, jobs_count: int = 32) -> dict:
    """Uneven multi-job schedule: jobs_count full registrations (steps=2 runs=2
    avg) with meshes of 10,242 / 40,962 / 163,842 (icosphere) or 61,442
    (Fibonacci hull) vertices and 128x64 / 256x128 / 512x256 maps, in one
    batched+adjoint schedule with tuned relaxation tiles.  Reports throughput,
    the per-job latency of each job run alone, and each job's largest vertex
    difference from the float64 plain path."""

    import numpy as np
    import torch

    from cat_surface_gpu import PipelineJob, run_cat_surface_gpu_pipelines, synthetic

    rng = np.random.default_rng(seed)
    meshes = (("icosphere", 10242, 5), ("icosphere", 40962, 6), ("fibonacci", 61442, 6),
              ("icosphere", 163842, 7))
    maps = ((128, 64), (256, 128), (512, 256))
    jobs, shapes = [], []
    t = time.perf_counter()
    for k in range(jobs_count):
        kind, n, sub = meshes[int(rng.integers(len(meshes)))]
        size = maps[int(rng.integers(len(maps)))]
        p = synthetic.make_pair(kind, n, subdivisions=sub, lmax=16, noise=0.05,
                                max_rotation_deg=30.0, warp_deg=3.0, seed=seed + 1000 + k)
        jobs.append(PipelineJob(meshes=_pair_meshes(p), map_size=size))
        shapes.append((n, size))
        log(f"uneven job {k + 1}/{jobs_count} built: {n} vertices, {size[0]}x{size[1]} map")
    log(f"{jobs_count} uneven jobs built in {time.perf_counter() - t:.1f} s")
    kw = dict(device=device, dartel_dtype="float32", steps=2, runs=2, avg=True,
              write_output=False, cuda_graph=True, rotation_method="adjoint")
    run_cat_surface_gpu_pipelines(jobs, autotune=True, **kw)  # capture graphs, tune tiles
    torch.cuda.synchronize()
    log("uneven warmup done")
    t0 = time.perf_counter()
    batched = run_cat_surface_gpu_pipelines(jobs, autotune=True, **kw)
    torch.cuda.synchronize()
    makespan = time.perf_counter() - t0
    log("uneven batched run done")
    # the same schedule as one captured graph: plan + capture, then a replay
    run_cat_surface_gpu_pipelines(jobs, autotune=True, graph_schedule=True, **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    graphed = run_cat_surface_gpu_pipelines(jobs, autotune=True, graph_schedule=True, **kw)
    torch.cuda.synchronize()
    makespan_graph = time.perf_counter() - t0
    replayed = all(r.timings.get("graph_replay") == 1.0 for r in graphed)
    same = all(np.array_equal(a.vertices, b.vertices) for a, b in zip(graphed, batched))
    log(f"uneven graph schedule: {makespan_graph:.3f} s (replayed {replayed}, equal {same})")
    alone, alone_graph, errors = [], [], []
    for job, res in zip(jobs, batched):
        run_cat_surface_gpu_pipelines([job], **kw)  # warm this job's graph shapes
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        run_cat_surface_gpu_pipelines([job], **kw)
        torch.cuda.synchronize()
        alone.append(time.perf_counter() - t1)
        run_cat_surface_gpu_pipelines([job], graph_schedule=True, **kw)  # plan + capture
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        run_cat_surface_gpu_pipelines([job], graph_schedule=True, **kw)
        torch.cuda.synchronize()
        alone_graph.append(time.perf_counter() - t1)
        ref = run_cat_surface_gpu_pipelines([job], **dict(kw, dartel_dtype="float64"))[0]
        errors.append(float(np.abs(res.vertices.astype(np.float64) - ref.vertices).max()))
        log(f"uneven job {len(alone)}/{len(jobs)}: alone {alone[-1]:.3f} s, graph "
            f"{alone_graph[-1]:.3f} s, max vertex delta vs float64 {errors[-1]:.2e}")
    lat = np.asarray(alone)
    lat_graph = np.asarray(alone_graph)
    return {
        "metric": "uneven multi-job CAT_SurfWarp throughput",
        "value": round(jobs_count / makespan, 3), "unit": "jobs/s", "n_gpus": 1,
        "higher_is_better": True, "dtype": "f32",
        "config": {"workload": f"{jobs_count} jobs, meshes 10,242-163,842 vertices, maps "
                               "128x64-512x256, steps=2 runs=2 avg, batched+adjoint, tuned "
                               "tiles", "jobs": [{"vertices": n, "map": f"{a}x{b}"}
                                                 for n, (a, b) in shapes]},
        "makespan_s": round(makespan, 4),
        "alone_latency_s": {"p50": round(float(np.percentile(lat, 50)), 4),
                            "p99": round(float(np.percentile(lat, 99)), 4),
                            "sum": round(float(lat.sum()), 4)},
        "graph_schedule": {
            "value": round(jobs_count / makespan_graph, 3), "unit": "jobs/s",
            "makespan_s": round(makespan_graph, 4), "replayed": replayed,
            "equal_to_multi_graph_schedule": same,
            "alone_latency_s": {"p50": round(float(np.percentile(lat_graph, 50)), 4),
                                "p99": round(float(np.percentile(lat_graph, 99)), 4),
                                "sum": round(float(lat_graph.sum()), 4)}},
        "batched_speedup_vs_alone_sum": round(float(lat.sum()) / makespan, 3),
        "max_vertex_delta_vs_float64_plain": {"max": max(errors),
                                              "p50": float(np.median(errors))},
    }


def hopf_report(seed: int, device) -> dict:
    """Rotation search on the 61,442-vertex LH/RH pair: the reference's 46 seeds
    + Nelder-Mead against 8,192 Hopf-grid candidates scored in one launch and
    the best 8 refined by 20 gradient steps each."""

    import torch

    from cat_surface_gpu import PipelineJob, run_cat_surface_gpu_pipelines

    pairs = make_inputs(seed)
    jobs = [PipelineJob(meshes=m) for m in pairs]
    kw = dict(device=device, dartel_dtype="float32", steps=2, runs=2, avg=True,
              write_output=False, cuda_graph=True)
    out = {}
    for name, extra in (("seeds+nelder-mead", {}),
                        ("hopf8192+adjoint", dict(rotation_method="adjoint",
                                                  rotation_candidates="hopf"))):
        run_cat_surface_gpu_pipelines(jobs, **kw, **extra)
        torch.cuda.synchronize()
        res = run_cat_surface_gpu_pipelines(jobs, **kw, **extra)
        out[name] = {"rotation_search_s": round(res[0].timings["rotation_search_seconds"], 4),
                     "costs": [r.rotation_cost for r in res],
                     "angles": [[round(float(a), 6) for a in r.rotation_angle] for r in res]}
    nm, hopf = out["seeds+nelder-mead"]["costs"], out["hopf8192+adjoint"]["costs"]
    return {"metric": "rotation search, 61,442-vertex LH+RH", "dtype": "f32 locate, f64 SSD",
            "n_gpus": 1, "results": out,
            "cost_ratio_hopf_over_nm": [h / c for h, c in zip(hopf, nm)],
            "within_1e-6_or_lower": all(h <= c * (1 + 1e-6) for h, c in zip(hopf, nm))}


def adjoint_report(device) -> dict:
    """The exp-warp-SSD adjoint: recompute vs store time and memory, float32
    against float64, and against the reference DARTEL code-0 rhs."""

    import torch

    from cat_surface_gpu.dartel_adjoint import (cosine, dartel_rhs, exp_warp_ssd,
                                                recompute_vs_store, relative_l2,
                                                synthetic_problem)

    rows = []
    for nx, ny in ((256, 128), (512, 256)):
        row = recompute_vs_store(nx, ny, device=device)
        v, f, t = synthetic_problem(nx, ny, device=device)
        grads = {}
        for dtype in (torch.float64, torch.float32):
            w = v.to(dtype).clone().requires_grad_(True)
            exp_warp_ssd(w, f.to(dtype), t.to(dtype)).backward()
            grads[dtype] = w.grad
        rhs = dartel_rhs(f, t, v)
        row["rel_l2_f32_vs_f64"] = relative_l2(grads[torch.float32], grads[torch.float64])
        row["rel_l2_vs_dartel_rhs"] = relative_l2(grads[torch.float32], rhs)
        row["cosine_vs_dartel_rhs"] = cosine(grads[torch.float32], rhs)
        rows.append(row)
    return {"metric": "exp-warp-SSD adjoint (K=6, membrane 0.01)", "n_gpus": 1, "rows": rows}


def sor_traffic_run(device, copies: int = 3) -> dict:
    """The dominant kernel alone, for hardware-counter runs: one depth-potential
    solve of four 40,962-vertex meshes (the bench's sides: icosphere-6
    resampling of each surface, 1000 SOR sweeps, the patch-blocked kernel), then
    `copies` plain device copies of 64 MiB (torch) whose counters give the
    bytes-per-counter-unit scale."""

    import numpy as np
    import torch

    import cat_surface_gpu as csg
    from cat_surface_gpu.rotation_feature import depth_potential
    from cat_surface_gpu.stencil_builder import icosphere

    v, f = icosphere(6)
    n = v.shape[0]
    idx = np.repeat(np.arange(n, dtype=np.int32)[:, None], 3, axis=1)
    w = np.tile(np.array([1.0, 0.0, 0.0]), (n, 1))
    st = csg.SurfaceStencil(sphere_points=v, faces=f.astype(np.int32), surface_indices=idx,
                            surface_weights=w, sheet_indices=np.zeros((1, 1, 3), np.int32),
                            sheet_weights=np.zeros((1, 1, 3)), nx=1, ny=1,
                            source_points=n).to(device)
    rng = np.random.default_rng(7)
    pts = torch.as_tensor(v[None] * 60.0 * (1.0 + 0.05 * rng.standard_normal((4, n, 1))),
                          dtype=torch.float32, device=device)
    depth_potential(st, pts)  # warm
    torch.cuda.synchronize()
    t = time.perf_counter()
    depth_potential(st, pts)
    torch.cuda.synchronize()
    solve_s = time.perf_counter() - t
    nbytes = 64 << 20
    src = torch.ones(nbytes // 4, dtype=torch.float32, device=device)
    dst = torch.empty_like(src)
    for _ in range(copies):
        torch.mul(src, 1.0, out=dst)  # a kernel copy (a device-to-device memcpy may bypass the shaders)
    torch.cuda.synchronize()
    return {"workload": "depth potential, 4 x 40,962 vertices, 1000 SOR sweeps",
            "solve_s": round(solve_s, 4), "copy_bytes": nbytes, "copies": copies}


def gifti_io_leg(pairs, args, device) -> dict:
    """The timed mode with GIFTI input and output: the four input surfaces of each
    hemisphere are written as GZipBase64Binary files (zlib level 6, nibabel's
    default) before timing; each timed step reads them, registers LH and RH and
    writes both output spheres, with the reader's caches cleared first."""

    import tempfile

    import torch

    from cat_surface_gpu import PipelineJob, run_cat_surface_gpu_pipelines, gifti, synthetic
    from cat_surface_gpu.synthetic import SyntheticPair

    folder = Path(tempfile.mkdtemp(prefix="csg_bench_"))
    jobs = []
    for k, (ss, sp, ts, tp) in enumerate(pairs):
        pair = SyntheticPair(sp.vertices, ss.vertices, sp.faces, tp.vertices, ts.vertices,
                             tp.faces, None)
        paths = synthetic.write_pair(pair, folder / f"h{k}", compress_level=6)
        jobs.append(PipelineJob(paths["src_surface"], paths["src_sphere"], paths["tgt_surface"],
                                paths["tgt_sphere"], folder / f"h{k}" / "out.sphere.reg.gii"))
    kw = dict(device=device, dartel_dtype=args.dtype, steps=2, runs=2, avg=True,
              write_output=True, cuda_graph=args.cuda_graph, autotune=True,
              rotation_method="adjoint")
    graph = args.mode == "graph" and args.cuda_graph

    def step() -> dict:
        if not graph:
            res = run_cat_surface_gpu_pipelines(jobs, **kw)
            return {k: res[0].timings[k] for k in ("input_read_seconds", "output_write_seconds")}
        # the graph schedule takes meshes in memory: read every file (in
        # parallel, as the pipeline reads them), replay, write
        from cat_surface_gpu.gpu_pipeline import _mesh_from_arrays, read_gifti_many

        t_read = time.perf_counter()
        paths = [p for j in jobs for p in (j.source_surface, j.source_sphere,
                                           j.target_surface, j.target_sphere)]
        arrays = read_gifti_many(paths)
        meshes = [_mesh_from_arrays(a, p) for a, p in zip(arrays, paths)]
        mem = [PipelineJob(output=j.output, meshes=tuple(meshes[4 * k:4 * k + 4]))
               for k, j in enumerate(jobs)]
        t_read = time.perf_counter() - t_read
        t_reg = time.perf_counter()
        run_cat_surface_gpu_pipelines(mem, graph_schedule=True, **kw)
        return {"input_read_seconds": t_read,
                "register_and_write_seconds": time.perf_counter() - t_reg}

    for _ in range(max(1, args.warmup)):
        step()
    torch.cuda.synchronize()
    sums: dict[str, float] = {}
    t = time.perf_counter()
    for _ in range(args.steps):
        gifti.clear_caches()
        for key, val in step().items():
            sums[key] = sums.get(key, 0.0) + val
    torch.cuda.synchronize()
    value = (time.perf_counter() - t) / (args.steps * len(jobs))
    out = {"value": round(value, 4), "unit": "s/hemisphere",
           "definition": "read the 4 GIFTI inputs of each hemisphere, register, write the "
                         "output sphere (GZipBase64Binary), LH+RH per step"
                         + (", registered by the captured graph" if graph else ""),
           "stages_s_per_hemisphere": {k: round(v / (args.steps * len(jobs)), 4)
                                       for k, v in sums.items()}}
    if graph:
        # the same work as a stream of batches: the next batch's files are read,
        # and the previous batch's outputs written, while the GPU registers
        from cat_surface_gpu import run_cat_surface_gpu_stream

        batches = [jobs] * max(4, args.steps)
        gifti.clear_caches()
        torch.cuda.synchronize()
        t = time.perf_counter()
        run_cat_surface_gpu_stream(batches, **kw)
        torch.cuda.synchronize()
        value = (time.perf_counter() - t) / (len(batches) * len(jobs))
        out["streamed"] = {
            "value": round(value, 4), "unit": "s/hemisphere", "batches": len(batches),
            "definition": "run_cat_surface_gpu_stream over batches of the LH+RH files: batch "
                          "k+1 read and batch k written while batch k registers; wall time of "
                          "the whole stream (first read and last write included) per hemisphere"}
    log("GIFTI leg done")
    return out


def instrumented_rooflines(pairs, args, device) -> list[dict]:
    """Per-kernel HIP-event timing of one step of the timed mode launched without
    graph capture, right after the timed region (a graph replay's kernels cannot
    be bracketed by host-side events): the same kernels on the same inputs."""

    import torch

    from cat_surface_gpu import _lib

    kernel_rooflines()  # clear
    torch.cuda.synchronize()
    _lib.lib().csg_kernel_timing(1)
    # the graph mode's kernels are the batched+adjoint schedule's
    run_step(pairs, args.dtype, "batched+adjoint" if args.mode == "graph" else args.mode,
             device, False)
    torch.cuda.synchronize()
    _lib.lib().csg_kernel_timing(0)
    return kernel_rooflines()


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--dtype", choices=("float32", "float64"), default="float32")
    ap.add_argument("--mode", choices=("plain", "plain-fused", "batched", "batched+adjoint",
                                       "graph"), default="graph")
    ap.add_argument("--seed", type=int, default=20261205)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-host-inputs", action="store_true",
                    help="skip the leg that hands the inputs over as host arrays each step")
    ap.add_argument("--no-gifti-io", action="store_true",
                    help="skip the leg that reads and writes GIFTI files around each step")
    ap.add_argument("--compare-plain", action=argparse.BooleanOptionalAction, default=True,
                    help="after the timed mode, time the plain path on the same inputs and "
                         "report it with the largest vertex difference between the two")
    ap.add_argument("--workload", choices=("hemisphere-pair", "uneven-32", "hopf-8192",
                                           "adjoint", "sor-traffic"), default="hemisphere-pair",
                    help="hemisphere-pair is the headline metric; the others print reports "
                         "of the uneven multi-job schedule, the Hopf-grid rotation search and "
                         "the exp-warp-SSD adjoint")
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
    if args.workload != "hemisphere-pair":
        report = {"uneven-32": lambda: uneven_report(args.seed, device),
                  "hopf-8192": lambda: hopf_report(args.seed, device),
                  "adjoint": lambda: adjoint_report(device),
                  "sor-traffic": lambda: sor_traffic_run(device)}[args.workload]()
        print(json.dumps(report), flush=True)
        return 0
    t = time.perf_counter()
    host_pairs = make_inputs(args.seed)
    pairs = resident_inputs(host_pairs, device)
    log(f"inputs ready in {time.perf_counter() - t:.1f} s")
    for i in range(args.warmup):
        run_step(pairs, args.dtype, args.mode, device, args.cuda_graph)
        torch.cuda.synchronize()
        log(f"warmup {i} done")
    torch.cuda.synchronize()
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
    ms_per_step = elapsed / args.steps * 1e3
    hemispheres = 2
    value = elapsed / (args.steps * hemispheres)
    kernels = instrumented_rooflines(pairs, args, device)
    log("instrumented step done")
    roof = dict(kernels[0], source="HIP events on the kernel's stream, one un-captured step "
                                   "of the timed mode after the timed region") if kernels else None
    plain = None
    if args.compare_plain and args.mode not in ("plain", "plain-fused"):
        plain = {}
        for leg in ("plain", "plain-fused"):
            for i in range(args.warmup):
                run_step(pairs, args.dtype, leg, device, args.cuda_graph)
            torch.cuda.synchronize()
            tp = time.perf_counter()
            for i in range(args.steps):
                plain_results = run_step(pairs, args.dtype, leg, device, args.cuda_graph)
            torch.cuda.synchronize()
            plain_value = (time.perf_counter() - tp) / (args.steps * hemispheres)
            delta = max(float(abs(a.vertices.astype("float64") - b.vertices).max())
                        for a, b in zip(results, plain_results))
            plain[leg] = {"value": round(plain_value, 4), "unit": "s/hemisphere",
                          "speedup": round(plain_value / value, 3),
                          "max_vertex_delta": delta}
            log(f"{leg} comparison done")
        plain["definition"] = PLAIN_DEFINITION
    host_leg = None if args.no_host_inputs else host_inputs_leg(host_pairs, args, device)
    gifti_leg = None if args.no_gifti_io else gifti_io_leg(host_pairs, args, device)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(host_pairs[0])
        cpu["gpu_speedup"] = round(cpu["value"] / value, 1)
        log("CPU baseline done")
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
                   "mode_definition": MODE_DEFINITIONS[args.mode],
                   "parallelism": "single GPU", "cuda_graph": args.cuda_graph},
        "plain": plain,
        "with_host_inputs": host_leg,
        "with_gifti_io": gifti_leg,
        "relax_tiles": _installed_tiles(),
        "roofline": roof,
        "roofline_kernels": kernels,
        "cpu_baseline": cpu,
        "stages_s_per_hemisphere": stages,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
