"""End-to-end CAT_SurfWarp-compatible registration on one MI355X.

Drop-in for run_cat_surface_gpu_pipeline (src/cat_surface_gpu/gpu_pipeline.py:
468-1326): same arguments, stages and result type.  What changes is where the
stages run: stencils are built in-process on the GPU (stencil_builder.py)
instead of by a CPU helper subprocess, rotation features come from the GPU
depth-potential solve, and all surfaces, stencils, maps and flows stay in HBM
between stages.  A stencil-builder or feature-helper *path* can still be
given, in which case that external program is run exactly as the reference
runs it.
"""

from __future__ import annotations

from dataclasses import dataclass
import math
import subprocess
import tempfile
import time
from pathlib import Path
from typing import Sequence

import numpy as np
import torch

from . import _lib
from .dartel_grid import resolve_device
from .gifti import read_gifti, write_gifti_surface, INTENT_POINTSET, INTENT_TRIANGLE
from .rotation_feature import compute_rotation_features
from .rotation_pipeline import RotationPipeline, read_rotation_values
from .stencil_builder import build_device_stencil
from .surface_rotation import rotation_matrix_np
from .surface_stencil import SurfaceStencil
from .surface_warp import apply_flow_to_stenciled_sphere, solve_dartel_from_surfaces


@dataclass(frozen=True)
class GiftiMesh:
    vertices: np.ndarray
    faces: np.ndarray


@dataclass(frozen=True)
class CatSurfaceGpuResult:
    vertices: np.ndarray
    faces: np.ndarray
    rotation_angle: np.ndarray
    rotation_cost: float
    timings: dict[str, float]
    output_path: Path


def read_gifti_mesh(path: str | Path) -> GiftiMesh:
    """First point-set and triangle arrays of a GIFTI surface (gpu_pipeline.py:52-71)."""

    arrays = read_gifti(path)
    if len(arrays) < 2:
        raise ValueError(f"GIFTI file must contain vertex and face arrays: {path}")
    vertices, faces = arrays[0].data, arrays[1].data
    if vertices.ndim != 2 or vertices.shape[1] != 3:
        raise ValueError(f"Vertex array must have shape [n, 3]: {path} -> {vertices.shape}")
    if faces.ndim != 2 or faces.shape[1] != 3:
        raise ValueError(f"Face array must have shape [m, 3]: {path} -> {faces.shape}")
    return GiftiMesh(np.ascontiguousarray(vertices, dtype=np.float32),
                     np.ascontiguousarray(faces, dtype=np.int32))


def write_gifti_mesh(path: str | Path, mesh: GiftiMesh, *,
                     reference_path: str | Path | None = None) -> None:
    """Write vertices and faces, keeping the reference file's intents (:74-110)."""

    point_intent, triangle_intent = INTENT_POINTSET, INTENT_TRIANGLE
    if reference_path is not None:
        arrays = read_gifti(reference_path)
        if len(arrays) >= 1:
            point_intent = arrays[0].intent
        if len(arrays) >= 2:
            triangle_intent = arrays[1].intent
    write_gifti_surface(path, mesh.vertices, mesh.faces, point_intent=point_intent,
                        triangle_intent=triangle_intent)


def _normalise_sphere(vertices: np.ndarray) -> np.ndarray:
    """Sequential float64 centroid, per-vertex float32 unit vectors (:274-305)."""

    value = np.ascontiguousarray(np.asarray(vertices, dtype=np.float32))
    if value.ndim != 2 or value.shape[1] != 3:
        raise ValueError(f"Sphere vertices must have shape [n, 3], got {value.shape}")
    centre = np.cumsum(value.astype(np.float64), axis=0)[-1]
    centred = (value.astype(np.float64) - centre / value.shape[0]).astype(np.float32)
    c64 = centred.astype(np.float64)
    length = np.sqrt((c64 * c64).sum(1))
    if not np.all(np.isfinite(length)) or np.any(length == 0.0):
        raise ValueError("Sphere contains a vertex that cannot be normalized")
    return np.ascontiguousarray((c64 * (1.0 / length)[:, None]).astype(np.float32))


def _rotation_matrix_from_angles(angles) -> np.ndarray:
    return rotation_matrix_np(angles)


def _rotate_vertices(vertices: np.ndarray, angles) -> np.ndarray:
    """p · Mᵀ in float64, stored float32 (:373-378)."""

    m = rotation_matrix_np(angles)
    return np.ascontiguousarray((np.asarray(vertices, np.float64) @ m.T).astype(np.float32))


def _resolve_dartel_dtype(value) -> torch.dtype:
    if isinstance(value, torch.dtype):
        resolved = value
    else:
        mapping = {"float32": torch.float32, "fp32": torch.float32, "float64": torch.float64,
                   "fp64": torch.float64}
        try:
            resolved = mapping[str(value).lower().replace("torch.", "")]
        except KeyError as error:
            raise ValueError("dartel_dtype must be float32/fp32 or float64/fp64") from error
    if resolved not in (torch.float32, torch.float64):
        raise ValueError("dartel_dtype must be torch.float32 or torch.float64")
    return resolved


def _run_checked(command: Sequence[str], label: str, environment=None) -> None:
    done = subprocess.run([str(c) for c in command], check=False, capture_output=True, text=True,
                          env=environment)
    if done.returncode != 0:
        raise RuntimeError(f"{label} failed with exit code {done.returncode}\n"
                           f"stdout:\n{done.stdout.strip()[-1000:]}\n"
                           f"stderr:\n{done.stderr.strip()[-2000:]}")


def _external_stencil(builder, sphere_path, out_path: Path, options: Sequence[str], device):
    _run_checked([str(builder), str(sphere_path), str(out_path), *options], "stencil builder")
    if not out_path.is_file() or out_path.stat().st_size == 0:
        raise RuntimeError(f"stencil builder did not produce a valid file: {out_path}")
    return SurfaceStencil.from_file(out_path).to(device)


def run_cat_surface_gpu_pipeline(
    source_surface_path, source_sphere_path, target_surface_path, target_sphere_path,
    output_path, *, source_stencil_path=None, target_stencil_path=None,
    rotation_values_path=None, rotation_values_probe=None, rotation_geometry_probe=None,
    rotation_depth_probe=None, stencil_builder=None, rotated_stencil_builder=None,
    stencil_threads: int = 8, device: str | torch.device = "cuda", kernel: str = "hip",
    dartel_dtype: str | torch.dtype = torch.float64, squaring_kernel: str | None = None,
    steps: int = 2, runs: int = 1, avg: bool = False, curvtypes: Sequence[int] = (5, 5, 2),
    fwhm: Sequence[float] = (5.0, 5.0 / 3.0, 5.0 / 9.0), loop: int = 6, cycles: int = 3,
    nit: int = 3, its: int = 3, code: int = 1, point_chunk: int = 4096,
    rotation_grid_size: int = 128, rotation_margin: int = 1, rotation_refine: bool = True,
    rotation_max_iter: int = 500, rotation_tol: float = 1.0e-4,
    rotation_simplex_step: float = 0.1, rotation_feature_backend: str = "auto",
    rotation_feature_solver: str = "colored-sor", parallel_sides: bool = True,
    optimized_dartel: bool = True, cuda_graph: bool = True,
    rotation_search=None, write_output: bool = True, meshes=None,
) -> CatSurfaceGpuResult:
    """Initial rotation, DARTEL runs, averaging, warp and GIFTI output (:468-1326).

    MI355X additions (keyword-only, defaults keep the reference behaviour):
    ``rotation_search`` is a callable (pipeline, source_points, source_values,
    target_values) -> RotationSearchResult replacing the 46-seed + Nelder-Mead
    search; ``meshes`` passes the four meshes (source surface, source sphere,
    target surface, target sphere) already in memory instead of reading the
    paths; ``write_output=False`` skips the output file.
    """

    if steps < 1 or steps > 3:
        raise ValueError("steps must be between 1 and 3")
    if runs < 1:
        raise ValueError("runs must be positive")
    if stencil_threads < 1 or stencil_threads > 256:
        raise ValueError("stencil_threads must be between 1 and 256")
    if len(curvtypes) < steps or len(fwhm) < steps:
        raise ValueError("curvtypes and fwhm must contain at least 'steps' entries")
    if rotation_max_iter < 1 or rotation_tol < 0.0 or rotation_simplex_step <= 0.0:
        raise ValueError("rotation_max_iter, rotation_tol, and rotation_simplex_step must be "
                         "positive")
    if rotation_feature_backend not in {"auto", "cuda", "cuda-official-geometry",
                                        "cuda-official-depth", "official-cpu"}:
        raise ValueError("rotation_feature_backend must be 'auto', 'cuda', "
                         "'cuda-official-geometry', 'cuda-official-depth' or 'official-cpu'")
    if rotation_feature_solver not in {"colored-sor", "pcg"}:
        raise ValueError("rotation_feature_solver must be 'colored-sor' or 'pcg'")
    solve_dtype = _resolve_dartel_dtype(dartel_dtype)
    if squaring_kernel not in {None, "auto", "hip", "torch", "triton"}:
        raise ValueError("squaring_kernel must be None, 'auto', or 'hip'")
    if rotation_feature_backend == "auto":
        rotation_feature_backend = "official-cpu" if rotation_values_probe is not None else "cuda"
    if device == "cpu" or str(device) == "cpu":
        raise ValueError("The explicit GPU pipeline requires a HIP device")
    dev = resolve_device(device)
    graph_requested = bool(cuda_graph and (runs + int(avg) >= 3))

    t_total = time.perf_counter()
    timings: dict[str, float] = {}
    if meshes is not None:
        src_surface, src_sphere, tgt_surface, tgt_sphere = meshes
    else:
        src_surface = read_gifti_mesh(source_surface_path)
        src_sphere = read_gifti_mesh(source_sphere_path)
        tgt_surface = read_gifti_mesh(target_surface_path)
        tgt_sphere = read_gifti_mesh(target_sphere_path)
    timings["input_read_seconds"] = time.perf_counter() - t_total

    with tempfile.TemporaryDirectory(prefix="cat_surface_gpu_") as tmp_name:
        tmp = Path(tmp_name)

        def stencil(sphere_vertices, sphere_path, surface, label, normalise, path=None,
                    builder=None):
            if path is not None:
                if not Path(path).is_file():
                    raise FileNotFoundError(f"{label} does not exist: {path}")
                return SurfaceStencil.from_file(path).to(dev)
            if builder is not None:
                if sphere_path is None:
                    sphere_path = tmp / f"{label}.gii"
                    write_gifti_surface(sphere_path, sphere_vertices, src_sphere.faces
                                        if label.startswith(("source", "rotated", "avg-source"))
                                        else tgt_sphere.faces)
                opts = ([] if normalise else ["--no-normalize"]) + [
                    "--threads", str(stencil_threads)]
                return _external_stencil(builder, sphere_path, tmp / f"{label}.stencil", opts,
                                         dev)
            faces = (src_sphere.faces if label.startswith(("source", "rotated", "avg-source"))
                     else tgt_sphere.faces)
            return build_device_stencil(sphere_vertices, faces, surface_vertices=surface,
                                        normalise=normalise, device=dev)

        t = time.perf_counter()
        src_stencil = stencil(src_sphere.vertices, source_sphere_path, src_surface.vertices,
                              "source", True, source_stencil_path, stencil_builder)
        tgt_stencil = stencil(tgt_sphere.vertices, target_sphere_path, tgt_surface.vertices,
                              "target", True, target_stencil_path, stencil_builder)
        torch.cuda.synchronize(dev)
        timings["initial_stencil_seconds"] = time.perf_counter() - t

        # rotation features (rotation_feature.py; gpu_pipeline.py:720-895)
        t = time.perf_counter()
        if rotation_values_path is not None:
            src_vals, tgt_vals = read_rotation_values(rotation_values_path)
        elif rotation_feature_backend == "official-cpu":
            if rotation_values_probe is None:
                raise ValueError("official-cpu rotation features require rotation_values_probe")
            out = tmp / "rotation-values.bin"
            _run_checked([rotation_values_probe, source_surface_path, source_sphere_path,
                          target_surface_path, target_sphere_path, out], "Rotation feature")
            src_vals, tgt_vals = read_rotation_values(out)
        else:
            depth = None
            if rotation_feature_backend == "cuda-official-depth":
                if rotation_depth_probe is None:
                    raise ValueError("cuda-official-depth requires rotation_depth_probe")
                import os

                env = os.environ.copy()
                env["CAT_SURFACE_GPU_ROTATION_RAW_DEPTH"] = "1"
                env["CAT_SURFACE_GPU_ROTATION_RAW_DEPTH_GEOMETRY"] = "1"
                out = tmp / "rotation-depth-values.bin"
                _run_checked([rotation_depth_probe, source_surface_path, source_sphere_path,
                              target_surface_path, target_sphere_path, out], "Rotation depth",
                             env)
                depth = read_rotation_values(out)
            elif rotation_feature_backend == "cuda-official-geometry":
                raise ValueError("cuda-official-geometry requires rotation_geometry_probe "
                                 "outputs, which this build does not produce")
            feats = compute_rotation_features(
                (src_stencil, tgt_stencil),
                (torch.as_tensor(src_surface.vertices, device=dev),
                 torch.as_tensor(tgt_surface.vertices, device=dev)),
                heat_fwhms=(15.0, 10.0),
                depth_values=None if depth is None else tuple(
                    torch.as_tensor(d, device=dev) for d in depth),
                solver=rotation_feature_solver)
            src_vals, tgt_vals = feats[0].values, feats[1].values
        torch.cuda.synchronize(dev)
        timings["rotation_feature_seconds"] = time.perf_counter() - t

        t = time.perf_counter()
        rot = RotationPipeline.from_stencil(tgt_stencil, device=dev, grid_size=rotation_grid_size,
                                            margin=rotation_margin)
        torch.cuda.synchronize(dev)
        timings["rotation_index_upload_seconds"] = time.perf_counter() - t

        t = time.perf_counter()
        src_points = src_stencil.sphere_points64 if src_stencil.sphere_points64 is not None \
            else src_stencil.sphere_points
        if rotation_search is not None:
            result = rotation_search(rot, src_points, src_vals, tgt_vals)
        else:
            result = rot.search(src_points, src_vals, tgt_vals, point_chunk=point_chunk,
                                refine=rotation_refine, max_iter=rotation_max_iter,
                                tol=rotation_tol, simplex_step=rotation_simplex_step)
        timings["rotation_search_seconds"] = time.perf_counter() - t
        angle = np.asarray(result.angle.detach().cpu(), dtype=np.float64).reshape(3)
        rotation_cost = float(result.cost.detach().cpu())

        current_sphere = _rotate_vertices(_normalise_sphere(src_sphere.vertices), angle)
        pole = np.asarray((0.0, np.pi / 2.0, 0.0), dtype=np.float64)
        avg_tgt_surface = _rotate_vertices(tgt_surface.vertices, pole)
        avg_tgt_sphere = _rotate_vertices(_normalise_sphere(tgt_sphere.vertices), pole)

        t = time.perf_counter()
        builder = rotated_stencil_builder
        cur_stencil = stencil(current_sphere, None, src_surface.vertices, "rotated-000", False,
                              builder=builder)
        avg_tgt_stencil = None
        if avg:
            avg_tgt_stencil = stencil(avg_tgt_sphere, None, tgt_surface.vertices, "avg-target",
                                      False, builder=builder)
        torch.cuda.synchronize(dev)
        timings["rotated_stencil_seconds"] = time.perf_counter() - t

        src_surface_dev = torch.as_tensor(src_surface.vertices, device=dev)
        tgt_surface_dev = torch.as_tensor(tgt_surface.vertices, device=dev)

        def solve_one(sv, tv, ss, ts, solve_fwhm):
            return solve_dartel_from_surfaces(
                sv, tv, ss, ts, fwhm=solve_fwhm, curvtypes=tuple(curvtypes), n_steps=steps,
                loop=loop, cycles=cycles, nit=nit, its=its, code=code, kernel=kernel,
                device=dev, dtype=solve_dtype, squaring_kernel=squaring_kernel,
                parallel_sides=parallel_sides, optimized=optimized_dartel,
                cuda_graph=graph_requested)

        t_dartel = time.perf_counter()
        last = None
        final_prewarp = None
        avg_src_stencil = avg_src_surface = avg_src_sphere = None
        for run in range(runs):
            t = time.perf_counter()
            run_fwhm = tuple(float(fwhm[0]) / (3.0 ** (steps * (run + 1) + i))
                             for i in range(steps))
            if avg and run == runs - 1:
                final_prewarp = current_sphere.copy()
                avg_src_surface = _rotate_vertices(src_surface.vertices, pole)
                avg_src_sphere = _rotate_vertices(final_prewarp, pole)
                avg_src_stencil = stencil(avg_src_sphere, None, src_surface.vertices,
                                          "avg-source", False, builder=builder)
            solve = solve_one(src_surface_dev, tgt_surface_dev, cur_stencil, tgt_stencil,
                              run_fwhm)
            if not (avg and run == runs - 1):
                warped = apply_flow_to_stenciled_sphere(current_sphere, solve.flow, cur_stencil,
                                                        inverse=True, device=dev,
                                                        dtype=solve_dtype)
                current_sphere = warped.to(torch.float32).cpu().numpy()
            last = solve
            if run + 1 < runs:
                cur_stencil = stencil(current_sphere, None, src_surface.vertices,
                                      f"rotated-{run + 1:03d}", False, builder=builder)
            torch.cuda.synchronize(dev)
            timings[f"dartel_run_{run}_seconds"] = time.perf_counter() - t
        timings["source_dartel_total_seconds"] = time.perf_counter() - t_dartel

        if avg:
            t = time.perf_counter()
            avg_fwhm = tuple(float(fwhm[0]) / (3.0 ** (steps * (runs + 1) + i))
                             for i in range(steps))
            avg_result = solve_one(torch.as_tensor(avg_src_surface, device=dev),
                                   torch.as_tensor(avg_tgt_surface, device=dev),
                                   avg_src_stencil, avg_tgt_stencil, avg_fwhm)
            main_warped = apply_flow_to_stenciled_sphere(final_prewarp, last.flow, cur_stencil,
                                                         inverse=True, device=dev,
                                                         dtype=solve_dtype)
            avg_warped = apply_flow_to_stenciled_sphere(avg_src_sphere, avg_result.flow,
                                                        avg_src_stencil, inverse=True,
                                                        device=dev, dtype=solve_dtype)
            unrotate = rotation_matrix_np(-pole)
            combined = torch.empty((main_warped.shape[0], 3), dtype=torch.float32, device=dev)
            _lib.check(_lib.lib().csg_average_xz(
                _lib.ptr(avg_warped), _lib.ptr(main_warped), int(main_warped.shape[0]),
                _lib.dtype_code(solve_dtype), _lib.doubles(unrotate.reshape(-1)),
                _lib.ptr(combined), _lib.stream_handle(dev)), "csg_average_xz")
            current_sphere = combined.cpu().numpy()
            timings["avg_seconds"] = time.perf_counter() - t
        timings["dartel_total_seconds"] = time.perf_counter() - t_dartel

        t = time.perf_counter()
        final = GiftiMesh(np.ascontiguousarray(current_sphere, dtype=np.float32),
                          src_sphere.faces)
        output = Path(output_path)
        if write_output:
            write_gifti_mesh(output, final,
                             reference_path=None if meshes is not None else source_sphere_path)
        timings["output_write_seconds"] = time.perf_counter() - t
    timings["total_seconds"] = time.perf_counter() - t_total
    return CatSurfaceGpuResult(vertices=final.vertices, faces=final.faces,
                               rotation_angle=angle, rotation_cost=rotation_cost,
                               timings=timings, output_path=output)
