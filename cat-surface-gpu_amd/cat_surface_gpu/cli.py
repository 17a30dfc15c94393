"""Command-line entry points ``cat-surfwarp-gpu`` and ``cat-surface-rotation-gpu``.

Same flags and JSON summary as src/cat_surface_gpu/cli.py:97-332 of the
reference.  Defaults that named the reference's bundled CPU helpers now
select the in-process GPU equivalents: stencils are built on the GPU unless
``--stencil-builder`` names an external program, and rotation features use
the GPU depth potential (``cuda`` backend) unless a helper is given.
"""

from __future__ import annotations

import argparse
import hashlib
import json
from pathlib import Path
import time

import numpy as np
import torch

from .dartel_grid import resolve_device
from .gpu_pipeline import run_cat_surface_gpu_pipeline
from .rotation_pipeline import RotationPipeline, read_rotation_points, read_rotation_values
from .surface_stencil import SurfaceStencil


def surfwarp_main(argv: list[str] | None = None) -> int:
    """Run the CAT_SurfWarp-compatible registration (cli.py:97-252)."""

    p = argparse.ArgumentParser(description="Run CAT_SurfWarp-compatible registration on MI355X.")
    p.add_argument("--source-surface", type=Path, required=True)
    p.add_argument("--source-sphere", type=Path, required=True)
    p.add_argument("--target-surface", type=Path, required=True)
    p.add_argument("--target-sphere", type=Path, required=True)
    p.add_argument("--output", type=Path, required=True)
    p.add_argument("--source-stencil", type=Path)
    p.add_argument("--target-stencil", type=Path)
    p.add_argument("--rotation-values", type=Path)
    p.add_argument("--rotation-values-probe", type=Path)
    p.add_argument("--rotation-geometry-probe", type=Path)
    p.add_argument("--rotation-depth-probe", type=Path)
    p.add_argument("--rotation-feature-backend",
                   choices=("auto", "cuda", "cuda-official-geometry", "cuda-official-depth",
                            "official-cpu"), default="auto")
    p.add_argument("--rotation-feature-solver", choices=("colored-sor", "pcg"),
                   default="colored-sor")
    p.add_argument("--stencil-builder", type=Path)
    p.add_argument("--rotated-stencil-builder", type=Path)
    p.add_argument("--stencil-threads", type=int, default=8)
    p.add_argument("--device", default="cuda")
    p.add_argument("--kernel", default="hip")
    p.add_argument("--dartel-dtype", choices=("float64", "float32"), default="float64")
    p.add_argument("--squaring-kernel", choices=("auto", "hip"), default="hip")
    p.add_argument("--steps", type=int, default=2)
    p.add_argument("--runs", type=int, default=2)
    p.add_argument("--avg", action=argparse.BooleanOptionalAction, default=True)
    p.add_argument("--loop", type=int, default=6)
    p.add_argument("--cycles", type=int, default=3)
    p.add_argument("--nit", type=int, default=3)
    p.add_argument("--its", type=int, default=3)
    p.add_argument("--code", type=int, default=1)
    p.add_argument("--parallel-sides", action=argparse.BooleanOptionalAction, default=True)
    p.add_argument("--optimized-dartel", action=argparse.BooleanOptionalAction, default=True)
    p.add_argument("--cuda-graph", action=argparse.BooleanOptionalAction, default=True)
    p.add_argument("--point-chunk", type=int, default=4096)
    p.add_argument("--rotation-grid-size", type=int, default=128)
    p.add_argument("--rotation-margin", type=int, default=1)
    p.add_argument("--no-rotation-refine", action="store_true")
    p.add_argument("--rotation-max-iter", type=int, default=500)
    p.add_argument("--rotation-tol", type=float, default=1.0e-4)
    p.add_argument("--rotation-simplex-step", type=float, default=0.1)
    # MI355X additions: gradient refinement of the best rotation candidates
    p.add_argument("--rotation-method", choices=("nelder-mead", "adjoint"),
                   default="nelder-mead")
    p.add_argument("--rotation-candidates", choices=("seeds", "hopf"), default="seeds")
    p.add_argument("--rotation-top", type=int, default=8)
    p.add_argument("--rotation-steps", type=int, default=20)
    args = p.parse_args(argv)
    result = run_cat_surface_gpu_pipeline(
        args.source_surface, args.source_sphere, args.target_surface, args.target_sphere,
        args.output, source_stencil_path=args.source_stencil,
        target_stencil_path=args.target_stencil, rotation_values_path=args.rotation_values,
        rotation_values_probe=args.rotation_values_probe,
        rotation_geometry_probe=args.rotation_geometry_probe,
        rotation_depth_probe=args.rotation_depth_probe,
        rotation_feature_backend=args.rotation_feature_backend,
        rotation_feature_solver=args.rotation_feature_solver,
        stencil_builder=args.stencil_builder,
        rotated_stencil_builder=args.rotated_stencil_builder or args.stencil_builder,
        stencil_threads=args.stencil_threads, device=args.device, kernel=args.kernel,
        dartel_dtype=args.dartel_dtype, squaring_kernel=args.squaring_kernel, steps=args.steps,
        runs=args.runs, avg=args.avg, loop=args.loop, cycles=args.cycles, nit=args.nit,
        its=args.its, code=args.code, point_chunk=args.point_chunk,
        rotation_grid_size=args.rotation_grid_size, rotation_margin=args.rotation_margin,
        rotation_refine=not args.no_rotation_refine, rotation_max_iter=args.rotation_max_iter,
        rotation_tol=args.rotation_tol, rotation_simplex_step=args.rotation_simplex_step,
        parallel_sides=args.parallel_sides, optimized_dartel=args.optimized_dartel,
        cuda_graph=args.cuda_graph, rotation_method=args.rotation_method,
        rotation_candidates=args.rotation_candidates, rotation_top=args.rotation_top,
        rotation_steps=args.rotation_steps)
    device = torch.device(args.device)
    print(json.dumps({
        "backend": "hip_gfx950_cat_surface_gpu_pipeline",
        "device": str(device),
        "device_name": torch.cuda.get_device_name(device) if device.type == "cuda" else "CPU",
        "torch": torch.__version__,
        "kernel": args.kernel,
        "steps": args.steps,
        "runs": args.runs,
        "avg": args.avg,
        "source_shape": list(result.vertices.shape),
        "faces_shape": list(result.faces.shape),
        "rotation_angle": result.rotation_angle.tolist(),
        "rotation_cost": result.rotation_cost,
        "timings": result.timings,
        "output": str(result.output_path),
    }, ensure_ascii=False, indent=2))
    return 0


def rotation_main(argv: list[str] | None = None) -> int:
    """Initial-rotation search on prepared inputs (cli.py:255-332)."""

    p = argparse.ArgumentParser(description="Run CAT initial-rotation search.")
    p.add_argument("--source-points", type=Path, required=True)
    p.add_argument("--target-stencil", type=Path, required=True)
    p.add_argument("--rotation-values", type=Path, required=True)
    p.add_argument("--device", default="auto")
    p.add_argument("--grid-size", type=int, default=128)
    p.add_argument("--margin", type=int, default=1)
    p.add_argument("--point-chunk", type=int, default=4096)
    p.add_argument("--candidate-table-dtype", choices=("int32", "int64"), default="int32")
    p.add_argument("--no-refine", action="store_true")
    p.add_argument("--max-iter", type=int, default=500)
    p.add_argument("--tol", type=float, default=1.0e-4)
    p.add_argument("--simplex-step", type=float, default=0.1)
    args = p.parse_args(argv)
    device = resolve_device(args.device)
    source_points = read_rotation_points(args.source_points)
    source_values, target_values = read_rotation_values(args.rotation_values)
    target_stencil = SurfaceStencil.from_file(args.target_stencil)
    started = time.perf_counter()
    pipeline = RotationPipeline.from_stencil(target_stencil, device=device,
                                             grid_size=args.grid_size, margin=args.margin)
    torch.cuda.synchronize(device)
    build_seconds = time.perf_counter() - started
    started = time.perf_counter()
    result = pipeline.search(source_points, source_values, target_values,
                             point_chunk=args.point_chunk, refine=not args.no_refine,
                             max_iter=args.max_iter, tol=args.tol, simplex_step=args.simplex_step)
    torch.cuda.synchronize(device)
    seed_costs = result.seed_costs.detach().cpu().numpy()
    print(json.dumps({
        "feature_backend": "c_probe_rotation_values",
        "cost_backend": "hip",
        "device": str(device),
        "device_name": torch.cuda.get_device_name(device),
        "source_points": int(source_points.shape[0]),
        "target_points": int(target_stencil.sphere_points.shape[0]),
        "build_upload_seconds": build_seconds,
        "search_seconds": time.perf_counter() - started,
        "seed_cost_sha256": hashlib.sha256(np.ascontiguousarray(seed_costs).tobytes()).hexdigest(),
        "angle": result.angle.detach().cpu().tolist(),
        "cost": float(result.cost.detach().cpu()),
        "rotation_matrix": result.rotation_matrix.detach().cpu().tolist(),
        "iterations": result.iterations,
    }, ensure_ascii=False, indent=2))
    return 0


if __name__ == "__main__":
    raise SystemExit(surfwarp_main())
