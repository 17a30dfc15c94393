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

import dataclasses
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass
import ctypes
import math
import subprocess
import tempfile
import time
from pathlib import Path
from typing import Sequence

import numpy as np
import torch

from . import _lib, schedule
from .dartel_grid import resolve_device
from .gifti import (INTENT_POINTSET, INTENT_TRIANGLE, read_gifti, read_gifti_intents,
                    read_gifti_many, write_gifti_surface)
from .rotation_feature import compute_rotation_features
from .rotation_pipeline import RotationPipeline, read_rotation_values
from .locate import check_all
from .stencil_builder import build_device_stencil
from .surface_rotation import rotation_matrix_np, search_rotations
from .surface_stencil import SurfaceStencil
from .surface_warp import apply_flow_to_stenciled_sphere, solve_dartel_from_surfaces_jobs


@dataclass(frozen=True)
class GiftiMesh:
    """Vertices [n, 3] and faces [m, 3]: host arrays, or tensors already on the
    device (then nothing is copied from the host, and a result's faces are the
    source sphere's tensor as given)."""

    vertices: np.ndarray
    faces: np.ndarray


def _host_array(a) -> np.ndarray:
    """A host copy of a device tensor; host arrays unchanged."""

    if isinstance(a, torch.Tensor):
        return a.detach().cpu().numpy()
    return np.asarray(a)


try:
    import xxhash

    def _content_hash(buf) -> int:
        return xxhash.xxh3_128_intdigest(buf)
except ImportError:  # the 128-bit BLAKE2 digest instead (~10x slower)
    import hashlib

    def _content_hash(buf) -> int:
        return int.from_bytes(hashlib.blake2b(buf, digest_size=16).digest(), "little")


def _faces_key(faces) -> tuple:
    """Identity of a face array's content: a 128-bit hash of a host array's bytes
    (0.15 ms for 122,880 triangles; the graph schedule keys every call on its
    face arrays); for a tensor, its storage, shape and version counter
    (in-place writes bump it)."""

    if isinstance(faces, torch.Tensor):
        return ("tensor", str(faces.device), faces.data_ptr(), tuple(faces.shape),
                faces._version)
    arr = np.ascontiguousarray(faces)
    return (arr.shape, arr.dtype.str, _content_hash(memoryview(arr).cast("B")))


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

    return _mesh_from_arrays(read_gifti(path), path)


def _mesh_from_arrays(arrays, path) -> GiftiMesh:
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
        intents = read_gifti_intents(reference_path)
        if len(intents) >= 1:
            point_intent = intents[0]
        if len(intents) >= 2:
            triangle_intent = intents[1]
    write_gifti_surface(path, _host_array(mesh.vertices), _host_array(mesh.faces),
                        point_intent=point_intent, triangle_intent=triangle_intent)


def _sphere_centroids(spheres: Sequence[torch.Tensor]) -> torch.Tensor:
    """[len(spheres), 3] float64 centroids, each the strictly sequential float64 sum
    of the vertices divided by n (the centring of gpu_pipeline.py:274-305), all in
    one launch (csg_sphere_centroids).  Asynchronous."""

    vs = [v.contiguous() for v in spheres]
    for v in vs:
        if v.ndim != 2 or v.shape[1] != 3 or v.dtype != torch.float32 or v.shape[0] < 1:
            raise ValueError("Sphere vertices must be a non-empty float32 [n, 3] tensor")
    dev = vs[0].device
    out = torch.empty((len(vs), 3), dtype=torch.float64, device=dev)
    ptrs = (ctypes.c_void_p * len(vs))(*[v.data_ptr() for v in vs])
    counts = (ctypes.c_int * len(vs))(*[int(v.shape[0]) for v in vs])
    _lib.check(_lib.lib().csg_sphere_centroids(ptrs, counts, len(vs), _lib.ptr(out),
                                               _lib.stream_handle(dev)), "csg_sphere_centroids")
    return out


def _prepare_sphere(vertices: torch.Tensor, *, centroid: torch.Tensor | None = None, angles=None,
                    bad: torch.Tensor | None = None) -> torch.Tensor:
    """The pipeline's sphere preparation (gpu_pipeline.py:274-305, :373-378) on the
    device (csg_prepare_sphere): with ``centroid`` (from _sphere_centroids), centre
    the vertices and scale each to unit length in float32 (a vertex that cannot be
    normalised sets ``bad``); with ``angles`` (host, or a device tensor whose
    matrix is then built on the device), rotate by p · Mᵀ in float64, stored
    float32.  Asynchronous."""

    if vertices.ndim != 2 or vertices.shape[1] != 3:
        raise ValueError(f"Sphere vertices must have shape [n, 3], got {tuple(vertices.shape)}")
    if centroid is not None and bad is None:
        raise ValueError("normalisation needs a device flag for bad vertices")
    v = vertices.to(torch.float32).contiguous()
    out = torch.empty_like(v)
    if isinstance(angles, torch.Tensor) and angles.device.type != "cpu":
        # the angles stay on the device (a rotation search's result): the
        # matrix is built there too
        ang = angles.to(torch.float64).reshape(1, 3).contiguous()
        mat = torch.empty(9, dtype=torch.float64, device=v.device)
        _lib.check(_lib.lib().csg_rotation_matrices(_lib.ptr(ang), 1, _lib.ptr(mat),
                                                    _lib.stream_handle(v.device)),
                   "csg_rotation_matrices")
        _lib.check(_lib.lib().csg_prepare_sphere_dev(_lib.ptr(v), int(v.shape[0]),
                                                     _lib.ptr(centroid), _lib.ptr(mat),
                                                     _lib.ptr(out), _lib.ptr(bad),
                                                     _lib.stream_handle(v.device)),
                   "csg_prepare_sphere_dev")
        return out
    rot = None if angles is None else _lib.doubles(rotation_matrix_np(angles).reshape(-1))
    _lib.check(_lib.lib().csg_prepare_sphere(_lib.ptr(v), int(v.shape[0]), _lib.ptr(centroid),
                                             rot, _lib.ptr(out), _lib.ptr(bad),
                                             _lib.stream_handle(v.device)),
               "csg_prepare_sphere")
    return out


def _sync(dev) -> None:
    """Wait for the device (stage timing), except while a schedule is captured."""

    if not schedule.capturing():
        torch.cuda.synchronize(dev)


_FACES_DEVICE: dict = {}


def _faces_on_device(faces, dev) -> torch.Tensor:
    """A face array on the device, cached per array (a captured schedule copies
    nothing from the host: the array must have been seen before)."""

    if isinstance(faces, torch.Tensor):
        return faces.to(device=dev)
    key = (str(dev),) + _faces_key(faces)
    hit = _FACES_DEVICE.get(key)
    if hit is None:
        if schedule.capturing():
            raise RuntimeError("face array not seen by an eager run of this schedule")
        hit = torch.as_tensor(np.ascontiguousarray(faces), device=dev)
        _FACES_DEVICE[key] = hit
    return hit


def _check_normalised(bad: torch.Tensor) -> None:
    if int(bad.item()):
        raise ValueError("Sphere contains a vertex that cannot be normalized")


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


@dataclass
class PipelineJob:
    """One registration of run_cat_surface_gpu_pipelines: the four input surfaces
    (GIFTI paths, or ``meshes`` = (source surface, source sphere, target
    surface, target sphere) already in memory), the output path and optional
    per-job stencil / rotation-value files."""

    source_surface: str | Path | None = None
    source_sphere: str | Path | None = None
    target_surface: str | Path | None = None
    target_sphere: str | Path | None = None
    output: str | Path | None = None
    meshes: tuple | None = None
    source_stencil: str | Path | None = None
    target_stencil: str | Path | None = None
    rotation_values: str | Path | None = None
    map_size: tuple[int, int] | None = None  # (nx, ny) of the DARTEL maps; None: 512 x 256


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
    rotation_method: str = "nelder-mead", rotation_candidates: str = "seeds",
    rotation_top: int = 8, rotation_steps: int = 20,
) -> CatSurfaceGpuResult:
    """Initial rotation, DARTEL runs, averaging, warp and GIFTI output (:468-1326).

    MI355X additions (keyword-only, defaults keep the reference behaviour):
    ``rotation_search`` is a callable (pipeline, source_points, source_values,
    target_values) -> RotationSearchResult replacing the 46-seed + Nelder-Mead
    search; ``meshes`` passes the four meshes (source surface, source sphere,
    target surface, target sphere) already in memory instead of reading the
    paths; ``write_output=False`` skips the output file.  ``rotation_method=
    "adjoint"`` refines the ``rotation_top`` best of the seeds (or of the
    8,192-rotation Hopf grid, ``rotation_candidates="hopf"``) by
    ``rotation_steps`` gradient steps instead of Nelder-Mead
    (surface_rotation.search_rotations).
    """

    job = PipelineJob(source_surface_path, source_sphere_path, target_surface_path,
                      target_sphere_path, output_path, meshes, source_stencil_path,
                      target_stencil_path, rotation_values_path)
    return run_cat_surface_gpu_pipelines(
        [job], rotation_values_probe=rotation_values_probe,
        rotation_geometry_probe=rotation_geometry_probe,
        rotation_depth_probe=rotation_depth_probe, stencil_builder=stencil_builder,
        rotated_stencil_builder=rotated_stencil_builder, stencil_threads=stencil_threads,
        device=device, kernel=kernel, dartel_dtype=dartel_dtype,
        squaring_kernel=squaring_kernel, steps=steps, runs=runs, avg=avg, curvtypes=curvtypes,
        fwhm=fwhm, loop=loop, cycles=cycles, nit=nit, its=its, code=code,
        point_chunk=point_chunk, rotation_grid_size=rotation_grid_size,
        rotation_margin=rotation_margin, rotation_refine=rotation_refine,
        rotation_max_iter=rotation_max_iter, rotation_tol=rotation_tol,
        rotation_simplex_step=rotation_simplex_step,
        rotation_feature_backend=rotation_feature_backend,
        rotation_feature_solver=rotation_feature_solver, parallel_sides=parallel_sides,
        optimized_dartel=optimized_dartel, cuda_graph=cuda_graph,
        rotation_search=rotation_search, write_output=write_output,
        rotation_method=rotation_method, rotation_candidates=rotation_candidates,
        rotation_top=rotation_top, rotation_steps=rotation_steps)[0]


@dataclass
class _JobState:
    job: PipelineJob
    src_surface: GiftiMesh
    src_sphere: GiftiMesh
    tgt_surface: GiftiMesh
    tgt_sphere: GiftiMesh
    timings: dict
    src_stencil: object = None
    tgt_stencil: object = None
    src_vals: object = None
    tgt_vals: object = None
    angle: np.ndarray | None = None
    rotation_cost: float = 0.0
    current_sphere: np.ndarray | None = None
    cur_stencil: object = None
    avg_tgt_surface: object = None  # device, float32
    avg_tgt_stencil: object = None
    avg_src_surface: object = None  # device, float32
    avg_src_sphere: object = None  # device, float32
    avg_src_stencil: object = None
    final_prewarp: object = None  # device, float32
    last_flow: object = None
    avg_flow: object = None
    src_dev: object = None
    tgt_dev: object = None
    src_faces_dev: object = None
    tgt_faces_dev: object = None
    current_dev: object = None  # device copy of current_sphere (float32)
    src_sphere_dev: object = None
    tgt_sphere_dev: object = None
    src_centroid: object = None  # device float64 [3]
    tgt_centroid: object = None

    def current_host(self) -> np.ndarray:
        if self.current_sphere is None:
            self.current_sphere = self.current_dev.cpu().numpy()
        return self.current_sphere


def run_cat_surface_gpu_pipelines(
    jobs: Sequence[PipelineJob], *, rotation_values_probe=None, rotation_geometry_probe=None,
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
    rotation_search=None, write_output: bool = True,
    rotation_method: str = "nelder-mead", rotation_candidates: str = "seeds",
    rotation_top: int = 8, rotation_steps: int = 20, autotune: bool = False,
    graph_schedule: bool = False,
) -> list[CatSurfaceGpuResult]:
    """Several independent CAT_SurfWarp registrations, stage by stage.

    Each job follows run_cat_surface_gpu_pipeline exactly (same stages, same
    results); what is shared is the schedule: the rotation features of every
    job and side are solved together, and every DARTEL round solves all jobs
    (and, in the last run, each job's main and average solves, which are
    independent: gpu_pipeline.py:1133-1224) in one batched solve.  Stage
    timings are those of the shared stages (the same for every job).

    ``graph_schedule=True`` runs the whole schedule (stencils, features,
    rotation search, every run's sheets and solves, warps, averaging) as one
    HIP graph: see run_graph_schedule.
    """

    if graph_schedule:
        kw = dict(locals())
        for name in ("jobs", "graph_schedule"):
            kw.pop(name)
        return run_graph_schedule(jobs, **kw)
    if not jobs:
        return []
    if steps < 1 or steps > 3:
        raise ValueError("steps must be between 1 and 3")
    if runs < 1:
        raise ValueError("runs must be positive")
    if stencil_threads < 1 or stencil_threads > 256:
        raise ValueError("stencil_threads must be between 1 and 256")
    if len(curvtypes) < steps or len(fwhm) < steps:
        raise ValueError("curvtypes and fwhm must contain at least 'steps' entries")
    for curvtype in curvtypes[:steps]:
        if int(curvtype) not in (2, 5):
            raise NotImplementedError("only upstream curvtype 2 and 5 exist")
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
    from .dartel_grid import _check_kernel

    _check_kernel(kernel)
    if squaring_kernel is not None:
        _check_kernel(squaring_kernel)
    if rotation_feature_backend == "auto":
        rotation_feature_backend = "official-cpu" if rotation_values_probe is not None else "cuda"
    if device == "cpu" or str(device) == "cpu":
        raise ValueError("The explicit GPU pipeline requires a HIP device")
    dev = resolve_device(device)
    graph_requested = bool(cuda_graph and (runs + int(avg) >= 3))
    pole = np.asarray((0.0, np.pi / 2.0, 0.0), dtype=np.float64)

    t_total = time.perf_counter()
    shared: dict[str, float] = {}
    states: list[_JobState] = []
    # every input file of every job read together (files and arrays in parallel threads)
    paths = [p for job in jobs if job.meshes is None
             for p in (job.source_surface, job.source_sphere, job.target_surface,
                       job.target_sphere)]
    arrays = iter(read_gifti_many(paths)) if paths else iter(())
    path_iter = iter(paths)
    for job in jobs:
        if job.meshes is not None:
            meshes = job.meshes
        else:
            meshes = tuple(_mesh_from_arrays(next(arrays), next(path_iter)) for _ in range(4))
        states.append(_JobState(job, *meshes, timings={}))
    shared["input_read_seconds"] = time.perf_counter() - t_total

    with tempfile.TemporaryDirectory(prefix="cat_surface_gpu_") as tmp_name:
        tmp = Path(tmp_name)

        def stencil(st: _JobState, index: int, sphere_vertices, sphere_path, label, normalise,
                    path=None, builder=None):
            source_side = label.startswith(("source", "rotated", "avg-source"))
            faces = st.src_sphere.faces if source_side else st.tgt_sphere.faces
            if path is not None:
                if not Path(path).is_file():
                    raise FileNotFoundError(f"{label} does not exist: {path}")
                return SurfaceStencil.from_file(path).to(dev)
            if builder is not None:
                if isinstance(sphere_vertices, torch.Tensor):
                    sphere_vertices = sphere_vertices.cpu().numpy()
                if sphere_path is None:
                    sphere_path = tmp / f"job{index}-{label}.gii"
                    write_gifti_surface(sphere_path, _host_array(sphere_vertices),
                                        _host_array(faces))
                opts = ([] if normalise else ["--no-normalize"]) + [
                    "--threads", str(stencil_threads)]
                return _external_stencil(builder, sphere_path,
                                         tmp / f"job{index}-{label}.stencil", opts, dev)
            return device_stencil(st, sphere_vertices, label, normalise)

        def device_stencil(st: _JobState, sphere_vertices, label, normalise, deferred=False):
            source_side = label.startswith(("source", "rotated", "avg-source"))
            surface = st.src_surface.vertices if source_side else st.tgt_surface.vertices
            faces_dev = st.src_faces_dev if source_side else st.tgt_faces_dev
            nx, ny = st.job.map_size or (512, 256)
            return build_device_stencil(sphere_vertices, faces_dev, surface_vertices=surface,
                                        normalise=normalise, nx=int(nx), ny=int(ny), device=dev,
                                        deferred=deferred)

        # In-process stencil builds of a stage are enqueued without host round trips
        # and confirmed together after the stage's synchronisation (settle); one
        # that fails its check (index capacity, search reach) is built again
        # synchronously before anything reads it.
        pending: list = []

        def build_into(st: _JobState, attr: str, index: int, sphere_vertices, label: str,
                       normalise: bool, builder, sphere_path=None, path=None):
            if builder is None and path is None:
                value, checks = device_stencil(st, sphere_vertices, label, normalise,
                                               deferred=True)
                pending.append((checks, st, attr, sphere_vertices, label, normalise))
            else:
                value = stencil(st, index, sphere_vertices, sphere_path, label, normalise, path,
                                builder)
            setattr(st, attr, value)

        def settle() -> None:
            if not pending:
                return
            if schedule.capturing():
                # the checks are read after the graph's replay
                for checks, *_ in pending:
                    for c in checks:
                        if c is not None:
                            schedule.defer(schedule._exceeds([c.status[0:1]], c.capacity),
                                           "stencil index capacity")
                            schedule.defer(c.status[1:2], "stencil query beyond the searched "
                                                          "reach")
                pending.clear()
                return
            verdict = check_all([c for entry in pending for c in entry[0]])
            k = 0
            for checks, st, attr, sphere_vertices, label, normalise in pending:
                ok = all(verdict[k:k + len(checks)])
                k += len(checks)
                if not ok:
                    setattr(st, attr, device_stencil(st, sphere_vertices, label, normalise))
            pending.clear()

        t = time.perf_counter()
        for i, st in enumerate(states):
            st.src_faces_dev = _faces_on_device(st.src_sphere.faces, dev)
            st.tgt_faces_dev = _faces_on_device(st.tgt_sphere.faces, dev)
            build_into(st, "src_stencil", i, st.src_sphere.vertices, "source", True,
                       stencil_builder, st.job.source_sphere, st.job.source_stencil)
            build_into(st, "tgt_stencil", i, st.tgt_sphere.vertices, "target", True,
                       stencil_builder, st.job.target_sphere, st.job.target_stencil)
            st.src_dev = torch.as_tensor(st.src_surface.vertices, device=dev)
            st.tgt_dev = torch.as_tensor(st.tgt_surface.vertices, device=dev)
            st.src_sphere_dev = torch.as_tensor(st.src_sphere.vertices, dtype=torch.float32,
                                                device=dev)
            st.tgt_sphere_dev = torch.as_tensor(st.tgt_sphere.vertices, dtype=torch.float32,
                                                device=dev)
        # every input sphere's centroid, for the normalisations after the rotation search
        centroids = _sphere_centroids([v for st in states
                                       for v in (st.src_sphere_dev, st.tgt_sphere_dev)])
        for i, st in enumerate(states):
            st.src_centroid, st.tgt_centroid = centroids[2 * i], centroids[2 * i + 1]
        _sync(dev)
        settle()
        shared["initial_stencil_seconds"] = time.perf_counter() - t

        # rotation features (rotation_feature.py; gpu_pipeline.py:720-895): every
        # job and side solved together where the topologies agree
        t = time.perf_counter()
        gpu_jobs = []
        depth = []
        for i, st in enumerate(states):
            job = st.job
            if job.rotation_values is not None:
                st.src_vals, st.tgt_vals = read_rotation_values(job.rotation_values)
                continue
            if rotation_feature_backend == "official-cpu":
                if rotation_values_probe is None:
                    raise ValueError("official-cpu rotation features require "
                                     "rotation_values_probe")
                out = tmp / f"job{i}-rotation-values.bin"
                _run_checked([rotation_values_probe, job.source_surface, job.source_sphere,
                              job.target_surface, job.target_sphere, out], "Rotation feature")
                st.src_vals, st.tgt_vals = read_rotation_values(out)
                continue
            if rotation_feature_backend == "cuda-official-depth":
                if rotation_depth_probe is None:
                    raise ValueError("cuda-official-depth requires rotation_depth_probe")
                import os

                env = os.environ.copy()
                env["CAT_SURFACE_GPU_ROTATION_RAW_DEPTH"] = "1"
                env["CAT_SURFACE_GPU_ROTATION_RAW_DEPTH_GEOMETRY"] = "1"
                out = tmp / f"job{i}-rotation-depth-values.bin"
                _run_checked([rotation_depth_probe, job.source_surface, job.source_sphere,
                              job.target_surface, job.target_sphere, out], "Rotation depth",
                             env)
                d = read_rotation_values(out)
                depth += [torch.as_tensor(d[0], device=dev), torch.as_tensor(d[1], device=dev)]
            elif rotation_feature_backend == "cuda-official-geometry":
                raise ValueError("cuda-official-geometry requires rotation_geometry_probe "
                                 "outputs, which this build does not produce")
            gpu_jobs.append(st)
        if gpu_jobs:
            feats = compute_rotation_features(
                [s for st in gpu_jobs for s in (st.src_stencil, st.tgt_stencil)],
                [v for st in gpu_jobs for v in (st.src_dev, st.tgt_dev)],
                heat_fwhms=tuple(f for _ in gpu_jobs for f in (15.0, 10.0)),
                depth_values=depth if depth else None, solver=rotation_feature_solver,
                parallel_sides=parallel_sides)
            for k, st in enumerate(gpu_jobs):
                st.src_vals, st.tgt_vals = feats[2 * k].values, feats[2 * k + 1].values
        _sync(dev)
        shared["rotation_feature_seconds"] = time.perf_counter() - t

        t = time.perf_counter()
        rots, problems = [], []
        for st in states:
            rot = RotationPipeline.from_stencil(st.tgt_stencil, device=dev,
                                                grid_size=rotation_grid_size,
                                                margin=rotation_margin)
            src_points = (st.src_stencil.sphere_points64
                          if st.src_stencil.sphere_points64 is not None
                          else st.src_stencil.sphere_points)
            rots.append((rot, src_points))
        _sync(dev)
        shared["rotation_index_upload_seconds"] = time.perf_counter() - t
        t = time.perf_counter()
        if rotation_search is not None:
            results = [rotation_search(rot, pts, st.src_vals, st.tgt_vals)
                       for (rot, pts), st in zip(rots, states)]
        else:
            # every job's seed scan and Nelder-Mead in lockstep (same results)
            problems = [(rot.index, rot.prepare_inputs(pts, st.src_vals, st.tgt_vals))
                        for (rot, pts), st in zip(rots, states)]
            if rotation_method == "adjoint" and rotation_candidates == "hopf":
                # the grid is scored on the features' maps (the stencil sheets)
                problems = [p + ((st.src_stencil.map_values_to_sheet(st.src_vals),
                                  st.tgt_stencil.map_values_to_sheet(st.tgt_vals)),)
                            for p, st in zip(problems, states)]
            results = search_rotations(problems, refine=rotation_refine,
                                       max_iter=rotation_max_iter, tol=rotation_tol,
                                       simplex_step=rotation_simplex_step,
                                       method=rotation_method,
                                       candidates=rotation_candidates, top=rotation_top,
                                       steps=rotation_steps)
        for st, result in zip(states, results):
            # kept on the device: the sphere's rotation matrix is built there
            st.angle = result.angle.detach().to(torch.float64).reshape(3)
            st.rotation_cost = result.cost.detach()
        shared["rotation_search_seconds"] = time.perf_counter() - t

        t = time.perf_counter()
        builder = rotated_stencil_builder
        # normalise and rotate on the device (gpu_pipeline.py:274-305, :373-378)
        bad = torch.zeros(1, dtype=torch.int32, device=dev)
        for i, st in enumerate(states):
            st.current_dev = _prepare_sphere(st.src_sphere_dev, centroid=st.src_centroid,
                                             angles=st.angle, bad=bad)
            st.current_sphere = None
            st.avg_tgt_surface = _prepare_sphere(st.tgt_dev, angles=pole)
            avg_tgt_sphere = _prepare_sphere(st.tgt_sphere_dev, centroid=st.tgt_centroid,
                                             angles=pole, bad=bad)
            build_into(st, "cur_stencil", i, st.current_dev, "rotated-000", False, builder)
            if avg:
                build_into(st, "avg_tgt_stencil", i, avg_tgt_sphere, "avg-target", False,
                           builder)
        _sync(dev)
        schedule.defer(bad, "Sphere contains a vertex that cannot be normalized", ValueError)
        settle()
        shared["rotated_stencil_seconds"] = time.perf_counter() - t

        # a recorded or captured schedule solves as the per-shape graphs do
        # (initial velocities taken as nonzero, no host check), and a captured
        # one launches the solves into the enclosing graph
        in_schedule = schedule.current() is not None
        solve_kw = dict(n_steps=steps, loop=loop, cycles=cycles, nit=nit, its=its, code=code,
                        curvtypes=tuple(int(t) for t in curvtypes[:steps]), device=dev,
                        dtype=solve_dtype,
                        cuda_graph=graph_requested and not schedule.capturing(),
                        assume_initial_nonzero=graph_requested or in_schedule)
        if autotune and not schedule.capturing():
            # relaxation tiles per map size, tuned (or read from the cache)
            # before any solve is captured; the batch is the largest solve's
            from .autotune import ensure_tuned

            t = time.perf_counter()
            sizes: dict[tuple[int, int], int] = {}
            for st in states:
                shape = (int(st.cur_stencil.nx), int(st.cur_stencil.ny))
                sizes[shape] = sizes.get(shape, 0) + steps * (2 if avg else 1)
            for shape, batch in sizes.items():
                ensure_tuned([shape], solve_dtype, batch, device=dev)
            shared["autotune_seconds"] = time.perf_counter() - t
        t_dartel = time.perf_counter()
        # each job's target side is the same surface and stencil in every run:
        # its fwhm-independent sheet preparation is made once
        prepared_sides: dict = {}
        for run in range(runs):
            t = time.perf_counter()
            run_fwhm = tuple(float(fwhm[0]) / (3.0 ** (steps * (run + 1) + i))
                             for i in range(steps))
            final_run = avg and run == runs - 1
            requests, owners = [], []
            for i, st in enumerate(states):
                if final_run:
                    st.final_prewarp = (st.current_dev.clone() if st.current_dev is not None
                                        else torch.as_tensor(st.current_host(), device=dev))
                    st.avg_src_surface = _prepare_sphere(st.src_dev, angles=pole)
                    st.avg_src_sphere = _prepare_sphere(st.final_prewarp, angles=pole)
                    build_into(st, "avg_src_stencil", i, st.avg_src_sphere, "avg-source",
                               False, builder)
                requests.append((st.src_dev, st.tgt_dev, st.cur_stencil, st.tgt_stencil,
                                 run_fwhm))
                owners.append((st, "main"))
            if final_run:
                # the average solve depends only on the pre-warp sphere of this run
                avg_fwhm = tuple(float(fwhm[0]) / (3.0 ** (steps * (runs + 1) + i))
                                 for i in range(steps))
                for st in states:
                    requests.append((torch.as_tensor(st.avg_src_surface, device=dev),
                                     torch.as_tensor(st.avg_tgt_surface, device=dev),
                                     st.avg_src_stencil, st.avg_tgt_stencil, avg_fwhm))
                    owners.append((st, "avg"))
            settle()  # the average solves read the avg-source stencils
            results = solve_dartel_from_surfaces_jobs(requests, prepared_sides=prepared_sides,
                                                      **solve_kw)
            for (st, role), res in zip(owners, results):
                if role == "avg":
                    st.avg_flow = res.flow
                    continue
                st.last_flow = res.flow
                if not final_run:
                    warped = apply_flow_to_stenciled_sphere(
                        st.current_host() if st.current_dev is None else st.current_dev,
                        res.flow, st.cur_stencil, inverse=True, device=dev, dtype=solve_dtype)
                    st.current_dev = warped.to(torch.float32)
                    st.current_sphere = None
            if run + 1 < runs:
                for i, st in enumerate(states):
                    sphere = (st.current_host() if builder is not None or st.current_dev is None
                              else st.current_dev)
                    build_into(st, "cur_stencil", i, sphere, f"rotated-{run + 1:03d}", False,
                               builder)
            _sync(dev)
            settle()
            shared[f"dartel_run_{run}_seconds"] = time.perf_counter() - t
        shared["source_dartel_total_seconds"] = time.perf_counter() - t_dartel

        if avg:
            t = time.perf_counter()
            unrotate = rotation_matrix_np(-pole)
            combined_all = []
            for st in states:
                main_warped = apply_flow_to_stenciled_sphere(st.final_prewarp, st.last_flow,
                                                             st.cur_stencil, inverse=True,
                                                             device=dev, dtype=solve_dtype)
                avg_warped = apply_flow_to_stenciled_sphere(st.avg_src_sphere, st.avg_flow,
                                                            st.avg_src_stencil, inverse=True,
                                                            device=dev, dtype=solve_dtype)
                combined = torch.empty((main_warped.shape[0], 3), dtype=torch.float32,
                                       device=dev)
                _lib.check(_lib.lib().csg_average_xz(
                    _lib.ptr(avg_warped), _lib.ptr(main_warped), int(main_warped.shape[0]),
                    _lib.dtype_code(solve_dtype), _lib.doubles(unrotate.reshape(-1)),
                    _lib.ptr(combined), _lib.stream_handle(dev)), "csg_average_xz")
                combined_all.append(combined)
            # one host copy after every job's launches (no synchronisation between jobs)
            for st, combined in zip(states, combined_all):
                st.current_dev = combined
                st.current_sphere = None if schedule.capturing() else combined.cpu().numpy()
            shared["avg_seconds"] = time.perf_counter() - t
        shared["dartel_total_seconds"] = time.perf_counter() - t_dartel

        t = time.perf_counter()
        results_out = []
        for st in states:
            if schedule.capturing():
                # the graph's outputs stay on the device; the caller reads them
                # after each replay
                results_out.append((st, GiftiMesh(st.current_dev, st.src_sphere.faces), Path("")))
                continue
            final = GiftiMesh(np.ascontiguousarray(st.current_host(), dtype=np.float32),
                              st.src_sphere.faces)
            output = Path(st.job.output) if st.job.output is not None else Path("")
            if write_output:
                if st.job.output is None:
                    raise ValueError("an output path is required when write_output is set")
                write_gifti_mesh(output, final, reference_path=None if st.job.meshes is not None
                                 else st.job.source_sphere)
            results_out.append((st, final, output))
        shared["output_write_seconds"] = time.perf_counter() - t
    shared["total_seconds"] = time.perf_counter() - t_total
    out = []
    capturing = schedule.capturing()
    angles = costs = None
    if not capturing and states:
        host = torch.cat([torch.cat([torch.as_tensor(st.angle, dtype=torch.float64).reshape(3),
                                     torch.as_tensor(st.rotation_cost,
                                                     dtype=torch.float64).reshape(1)])
                          .to("cpu") for st in states]).numpy().reshape(-1, 4)
        angles, costs = host[:, :3], host[:, 3]
    for k, (st, final, output) in enumerate(results_out):
        timings = dict(shared)
        if capturing:
            angle, cost = st.angle, st.rotation_cost
        else:
            angle, cost = np.array(angles[k]), float(costs[k])
        out.append(CatSurfaceGpuResult(vertices=final.vertices, faces=final.faces,
                                       rotation_angle=angle, rotation_cost=cost,
                                       timings=timings, output_path=output))
    return out


# ----------------------------------------------------------------- one graph

@dataclass
class _GraphSchedule:
    graph: object                 # torch.cuda.CUDAGraph
    inputs: list                  # per job: its four vertex arrays (device float32)
    outputs: list                 # per job: CatSurfaceGpuResult with device tensors
    capture: object               # schedule.Capture (bounds and device checks)
    capture_seconds: float
    replays: int = 0
    fallbacks: int = 0
    staging: list | None = None   # pinned host copies of `inputs` (host meshes go through them)
    out_pins: list | None = None  # pinned host (vertices, angle, cost) per job


_GRAPH_SCHEDULES: dict = {}


def clear_graph_schedules() -> None:
    _GRAPH_SCHEDULES.clear()


def _graph_key(jobs, kw, dev) -> tuple:
    parts = []
    for job in jobs:
        meshes = tuple((tuple(np.shape(m.vertices)), _faces_key(m.faces))
                       for m in job.meshes)
        parts.append(meshes + (tuple(job.map_size or (512, 256)),))
    opts = tuple(sorted((k, repr(v)) for k, v in kw.items() if k != "write_output"))
    return (str(dev), tuple(parts), opts)


def run_graph_schedule(jobs: Sequence[PipelineJob], **kw) -> list[CatSurfaceGpuResult]:
    """run_cat_surface_gpu_pipelines as one HIP graph (north-star pattern (b)).

    The first call for a given set of job shapes (vertex counts, face arrays and
    map sizes) and options runs the schedule eagerly -- that is its result --
    while recording how many launches each data-dependent loop needed, then
    captures the same schedule into one graph on static input buffers: every
    stage's kernels and the batched solves of every map size.  Later calls copy
    their meshes into those buffers and replay the graph.  Loop counts that
    depend on the data (heat-kernel iterations) are read by the kernels from
    device memory, and the graph launches up to the recorded bound x 1.25 + 8;
    each replay then checks on the device, and reads once, that every count
    stayed within its bound and every stencil index, SOR barrier, sphere
    normalisation and rotation cover is valid -- otherwise the call runs
    eagerly instead (and counts a fallback).

    Supported: meshes given in memory, in-process stencils and rotation
    features (SOR depth), the adjoint rotation search, and steps >= 2 of type-5
    sheets (the batched rule of surface_warp.py:427-448).
    """

    if not jobs:
        return []
    for job in jobs:
        if (job.meshes is None or job.source_stencil is not None or
                job.target_stencil is not None or job.rotation_values is not None):
            raise ValueError("the graph schedule needs in-memory meshes and in-process stencils "
                             "and rotation features")
    unsupported = [name for name in ("rotation_values_probe", "rotation_geometry_probe",
                                     "rotation_depth_probe", "stencil_builder",
                                     "rotated_stencil_builder", "rotation_search")
                   if kw.get(name) is not None]
    if unsupported:
        raise ValueError(f"the graph schedule does not run {', '.join(unsupported)}")
    if kw.get("rotation_method", "nelder-mead") != "adjoint":
        raise ValueError("the graph schedule needs rotation_method='adjoint' (Nelder-Mead "
                         "decides each step on the host)")
    if kw.get("rotation_feature_backend", "auto") not in ("auto", "cuda"):
        raise ValueError("the graph schedule computes rotation features on the GPU ('cuda')")
    if kw.get("rotation_feature_solver", "colored-sor") != "colored-sor":
        raise ValueError("the graph schedule solves the depth potential by SOR")
    steps = int(kw.get("steps", 2))
    types = tuple(int(t) for t in kw.get("curvtypes", (5, 5, 2))[:steps])
    if steps < 2 or any(t != 5 for t in types):
        raise ValueError("the graph schedule needs steps >= 2 of type-5 curvature")
    dev = resolve_device(kw.get("device", "cuda"))
    write_output = bool(kw.get("write_output", True))
    eager_kw = dict(kw, write_output=False)
    key = _graph_key(jobs, kw, dev)
    entry = _GRAPH_SCHEDULES.get(key)
    if entry is None:
        with schedule.recording() as rec:
            results = run_cat_surface_gpu_pipelines(jobs, **eager_kw)
        _GRAPH_SCHEDULES[key] = _capture_schedule(jobs, eager_kw, rec.bounds, dev)
        return _finish_graph_results(jobs, results, write_output)
    t0 = time.perf_counter()
    if entry.staging is None:
        entry.staging = [[torch.empty(b.shape, dtype=torch.float32, pin_memory=True)
                          for b in bufs] for bufs in entry.inputs]
        entry.out_pins = [tuple(torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
                                for t in (o.vertices, o.rotation_angle, o.rotation_cost))
                          for o in entry.outputs]
    # host meshes through pinned buffers, asynchronously on the replay's stream
    # (the buffers are free again: every call ends with a synchronising read)
    for job, bufs, pins in zip(jobs, entry.inputs, entry.staging):
        for mesh, buf, pin in zip(job.meshes, bufs, pins):
            if isinstance(mesh.vertices, torch.Tensor):
                buf.copy_(mesh.vertices.to(torch.float32), non_blocking=False)
            else:
                np.copyto(pin.numpy(), np.asarray(mesh.vertices), casting="same_kind")
                buf.copy_(pin, non_blocking=True)
    entry.graph.replay()
    # the outputs' host copies are queued before the checks' read, which waits for them
    for o, pins in zip(entry.outputs, entry.out_pins):
        for pin, t in zip(pins, (o.vertices, o.rotation_angle, o.rotation_cost)):
            pin.copy_(t, non_blocking=True)
    failed = entry.capture.verdict()
    entry.replays += 1
    if failed:
        entry.fallbacks += 1
        with schedule.recording():
            results = run_cat_surface_gpu_pipelines(jobs, **eager_kw)
        for r in results:
            r.timings["graph_fallback"] = 1.0
        return _finish_graph_results(jobs, results, write_output)
    if not entry.capture.flags:
        torch.cuda.current_stream(dev).synchronize()
    host = [(pv.numpy().copy(), np.asarray(pa.numpy(), dtype=np.float64).copy(), float(pc))
            for pv, pa, pc in entry.out_pins]
    total = time.perf_counter() - t0
    results = []
    for job, o, (verts, angle, cost) in zip(jobs, entry.outputs, host):
        results.append(CatSurfaceGpuResult(
            vertices=np.ascontiguousarray(verts, dtype=np.float32),
            faces=(job.meshes[1].faces if isinstance(job.meshes[1].faces, torch.Tensor)
                   else np.asarray(job.meshes[1].faces)),
            rotation_angle=angle, rotation_cost=cost,
            timings={"total_seconds": total, "graph_replay": 1.0},
            output_path=Path(job.output) if job.output is not None else Path("")))
    return _finish_graph_results(jobs, results, write_output)


def _read_job_meshes(batch: Sequence[PipelineJob]) -> list[PipelineJob]:
    """The batch with every job's four GIFTI inputs read into memory (all files
    of the batch in parallel threads); jobs given in memory pass unchanged."""

    paths = [p for j in batch if j.meshes is None
             for p in (j.source_surface, j.source_sphere, j.target_surface, j.target_sphere)]
    arrays = iter(read_gifti_many(paths)) if paths else iter(())
    path_iter = iter(paths)
    out = []
    for j in batch:
        if j.meshes is not None:
            out.append(j)
            continue
        meshes = tuple(_mesh_from_arrays(next(arrays), next(path_iter)) for _ in range(4))
        out.append(dataclasses.replace(j, meshes=meshes, source_surface=None,
                                       source_sphere=None, target_surface=None,
                                       target_sphere=None))
    return out


def run_cat_surface_gpu_stream(batches, **kw) -> list[list[CatSurfaceGpuResult]]:
    """Batches of registrations (GIFTI jobs, e.g. LH + RH of one subject per
    batch) through the captured schedule (run_graph_schedule, same options),
    one after another with the file work off the GPU's path: batch k + 1's
    inputs are read in a background thread while batch k registers, and batch
    k's output spheres are written while batch k + 1 registers.  Each batch's
    results equal run_cat_surface_gpu_pipelines(batch, graph_schedule=True);
    every output is written when the call returns."""

    write_output = bool(kw.pop("write_output", True))
    kw.pop("graph_schedule", None)
    batches = [list(b) for b in batches]
    if not batches:
        return []
    out: list[list[CatSurfaceGpuResult]] = []
    io = ThreadPoolExecutor(max_workers=2, thread_name_prefix="csg-stream")
    try:
        pending = []
        upcoming = io.submit(_read_job_meshes, batches[0])
        for k in range(len(batches)):
            jobs = upcoming.result()
            if k + 1 < len(batches):
                upcoming = io.submit(_read_job_meshes, batches[k + 1])
            results = run_cat_surface_gpu_pipelines(jobs, graph_schedule=True,
                                                    write_output=False, **kw)
            if write_output:
                pending.append(io.submit(_finish_graph_results, jobs, results, True))
            out.append(results)
        for f in pending:
            f.result()
    finally:
        io.shutdown(wait=True)
    return out


# output writers (threads start on first use)
_WRITE_POOL = ThreadPoolExecutor(max_workers=4, thread_name_prefix="csg-write")


def _finish_graph_results(jobs, results, write_output: bool):
    """Write each job's output sphere (the jobs' files in parallel threads)."""

    if write_output:
        if any(job.output is None for job in jobs):
            raise ValueError("an output path is required when write_output is set")
        list(_WRITE_POOL.map(
            lambda jr: write_gifti_mesh(Path(jr[0].output), GiftiMesh(jr[1].vertices, jr[1].faces)),
            zip(jobs, results)))
    return results


def _capture_schedule(jobs, kw, bounds, dev) -> _GraphSchedule:
    started = time.perf_counter()
    inputs = [[torch.as_tensor(m.vertices).to(device=dev, dtype=torch.float32).clone()
               for m in job.meshes] for job in jobs]
    cap_jobs = [PipelineJob(meshes=tuple(GiftiMesh(buf, m.faces)
                                         for buf, m in zip(bufs, job.meshes)),
                            map_size=job.map_size)
                for job, bufs in zip(jobs, inputs)]
    graph = torch.cuda.CUDAGraph()
    torch.cuda.synchronize(dev)
    current = torch.cuda.current_stream(dev)
    stream = torch.cuda.Stream(device=dev)
    stream.wait_stream(current)
    with torch.cuda.stream(stream), schedule.capture(bounds) as cap:
        with torch.cuda.graph(graph, stream=stream):
            outputs = run_cat_surface_gpu_pipelines(cap_jobs, **kw)
        if cap.position != len(cap.bounds):
            raise RuntimeError("the captured schedule issued fewer bounded loops than the "
                               "recorded one")
    current.wait_stream(stream)
    torch.cuda.synchronize(dev)
    return _GraphSchedule(graph=graph, inputs=inputs, outputs=outputs, capture=cap,
                          capture_seconds=time.perf_counter() - started)
