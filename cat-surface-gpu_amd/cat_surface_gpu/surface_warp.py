"""DARTEL solve, warp application and the solve cache, on MI355X.

Drop-in for the reference's src/cat_surface_gpu/surface_warp.py.  The whole
multi-step solve (``solve_dartel_maps``) is one call into libcsg_hip.so, which
launches every kernel of every Gauss-Newton step on the current stream; with
``cuda_graph=True`` that stream of launches is captured once per shape into a
hipGraph and replayed (the reference's graph cache, surface_warp.py:27-236).
"""

from __future__ import annotations

from dataclasses import dataclass
import time
from typing import Sequence

import torch

from . import _lib, schedule
from .dartel_grid import GridSpec, _check_kernel, resolve_device


@dataclass(frozen=True)
class DartelSolveResult:
    """Store a solved flow, warped feature maps, and DARTEL diagnostics."""

    flow: torch.Tensor
    source_maps: torch.Tensor
    target_maps: torch.Tensor
    metrics: torch.Tensor


def default_dartel_parameters(loop: int = 6, *, mu: float = 0.125, lambda_: float = 0.0,
                              lmreg: float = 1e-3, muchange: int = 4, murate: float = 1.25):
    """Upstream (1, 1, mu, lambda, lambda/2) per loop (surface_warp.py:76-107)."""

    if loop <= 0:
        raise ValueError(f"loop must be positive, got {loop}")
    if muchange <= 0 or murate <= 0:
        raise ValueError("muchange and murate must be positive")
    values = []
    current_mu = float(mu)
    current_lambda = float(lambda_)
    for index in range(loop):
        values.append((1.0, 1.0, current_mu, current_lambda, current_lambda / 2.0))
        if (index + 1) % muchange == 0:
            current_mu /= murate
        current_lambda /= 5.0
    return tuple(values)


@dataclass
class _SolveGraph:
    graph: torch.cuda.CUDAGraph
    source_maps: torch.Tensor
    target_maps: torch.Tensor
    flow: torch.Tensor
    metrics: torch.Tensor
    workspace: torch.Tensor
    capture_seconds: float


_DARTEL_CUDA_GRAPH_CACHE: dict[tuple, _SolveGraph] = {}


def _solve_key(source: torch.Tensor, target: torch.Tensor, **kw) -> tuple:
    return (source.device.type, source.device.index, source.dtype, tuple(source.shape),
            tuple(target.shape)) + tuple(sorted(kw.items()))


def _launch_solve(source: torch.Tensor, target: torch.Tensor, *, steps_per_job: int, loop: int,
                  lmreg: float, cycles: int, nit: int, its: int, code: int,
                  assume_initial_nonzero: bool, flow: torch.Tensor, metrics: torch.Tensor | None,
                  workspace: torch.Tensor, velocity: torch.Tensor | None = None) -> None:
    batch, ny, nx = (int(d) for d in source.shape)
    dcode = _lib.dtype_code(source.dtype)
    rc = _lib.lib().csg_solve_dartel_maps(
        _lib.ptr(source), _lib.ptr(target), nx, ny, batch, int(steps_per_job), int(loop),
        float(lmreg), int(cycles), int(nit), int(its), int(code),
        int(bool(assume_initial_nonzero)), dcode, _lib.ptr(flow), _lib.ptr(velocity),
        _lib.ptr(metrics), _lib.ptr(workspace), workspace.numel(),
        _lib.stream_handle(source.device))
    _lib.check(rc, "csg_solve_dartel_maps")


def solve_workspace(nx: int, ny: int, batch: int, dtype: torch.dtype,
                    device: torch.device) -> torch.Tensor:
    nbytes = _lib.lib().csg_solve_workspace_bytes(nx, ny, batch, _lib.dtype_code(dtype))
    return _lib.workspace(nbytes, device)


def _capture(source, target, *, loop, lmreg, cycles, nit, its, code, device,
             steps_per_job: int | None = None) -> _SolveGraph:
    static_source = source.detach().clone().contiguous()
    static_target = target.detach().clone().contiguous()
    steps, ny, nx = (int(d) for d in source.shape)
    per_job = steps if steps_per_job is None else int(steps_per_job)
    flow = torch.empty((steps // per_job, 2, ny, nx), dtype=source.dtype, device=device)
    metrics = torch.empty((steps, loop * its, 3), dtype=torch.float64, device=device)
    ws = solve_workspace(nx, ny, steps, source.dtype, device)
    graph = torch.cuda.CUDAGraph()
    torch.cuda.synchronize(device)
    started = time.perf_counter()
    current = torch.cuda.current_stream(device)
    capture_stream = torch.cuda.Stream(device=device)
    capture_stream.wait_stream(current)
    with torch.cuda.stream(capture_stream):
        with torch.cuda.graph(graph, stream=capture_stream):
            _launch_solve(static_source, static_target, steps_per_job=per_job, loop=loop,
                          lmreg=lmreg, cycles=cycles, nit=nit, its=its, code=code,
                          assume_initial_nonzero=True, flow=flow, metrics=metrics,
                          workspace=ws)
    current.wait_stream(capture_stream)
    torch.cuda.synchronize(device)
    return _SolveGraph(graph, static_source, static_target, flow, metrics, ws,
                       time.perf_counter() - started)


def prepare_dartel_cuda_graph(source_maps, target_maps, *, loop: int = 6, lmreg: float = 1e-3,
                              cycles: int = 3, nit: int = 3, its: int = 3, code: int = 1,
                              kernel: str = "auto", squaring_kernel: str | None = None,
                              device: str | torch.device = "auto",
                              dtype: torch.dtype = torch.float64, optimized: bool = True) -> float:
    """Capture the solve for this shape once (surface_warp.py:169-231)."""

    dev = resolve_device(device)
    _check_kernel(kernel)
    source = torch.as_tensor(source_maps, dtype=dtype, device=dev).contiguous()
    target = torch.as_tensor(target_maps, dtype=dtype, device=dev).contiguous()
    if tuple(source.shape) != tuple(target.shape):
        raise ValueError("source_maps and target_maps must have the same shape")
    key = _solve_key(source, target, loop=loop, lmreg=float(lmreg), cycles=cycles, nit=nit,
                     its=its, code=code)
    if key in _DARTEL_CUDA_GRAPH_CACHE:
        return 0.0
    entry = _capture(source, target, loop=loop, lmreg=lmreg, cycles=cycles, nit=nit, its=its,
                     code=code, device=dev)
    _DARTEL_CUDA_GRAPH_CACHE[key] = entry
    return entry.capture_seconds


def clear_dartel_cuda_graph_cache() -> None:
    _DARTEL_CUDA_GRAPH_CACHE.clear()


def solve_dartel_maps(source_maps, target_maps, *, loop: int = 6, lmreg: float = 1e-3,
                      cycles: int = 3, nit: int = 3, its: int = 3, code: int = 1,
                      kernel: str = "auto", squaring_kernel: str | None = None,
                      device: str | torch.device = "auto", dtype: torch.dtype = torch.float64,
                      optimized: bool = True, assume_initial_nonzero: bool = False,
                      cuda_graph: bool = False) -> DartelSolveResult:
    """Solve DARTEL feature maps with the upstream loop semantics (surface_warp.py:239-367).

    The ``steps`` maps are independent solves (velocity reset per step); they
    run batched in every kernel launch.  The flow is expdef(v of the last
    step, 10).
    """

    if cycles < 0 or nit < 0 or its < 0:
        raise ValueError("cycles, nit, and its must be non-negative")
    if squaring_kernel not in {None, "auto", "hip", "torch", "triton"}:
        raise ValueError("squaring_kernel must be None, 'auto', or 'hip'")
    _check_kernel(kernel)
    if squaring_kernel is not None:
        _check_kernel(squaring_kernel)
    dev = resolve_device(device)
    source = torch.as_tensor(source_maps, dtype=dtype, device=dev)
    target = torch.as_tensor(target_maps, dtype=dtype, device=dev)
    if source.ndim != 3 or target.ndim != 3:
        raise ValueError("source_maps and target_maps must have shape [steps, ny, nx]")
    if tuple(source.shape) != tuple(target.shape):
        raise ValueError("source_maps and target_maps must have the same shape")
    if source.shape[0] < 1:
        raise ValueError("At least one DARTEL step is required")
    if loop <= 0:
        raise ValueError(f"loop must be positive, got {loop}")
    source = source.contiguous()
    target = target.contiguous()
    steps, ny, nx = (int(d) for d in source.shape)
    if cuda_graph:
        key = _solve_key(source, target, loop=loop, lmreg=float(lmreg), cycles=cycles, nit=nit,
                         its=its, code=code)
        entry = _DARTEL_CUDA_GRAPH_CACHE.get(key)
        if entry is None:
            entry = _capture(source, target, loop=loop, lmreg=lmreg, cycles=cycles, nit=nit,
                             its=its, code=code, device=dev)
            _DARTEL_CUDA_GRAPH_CACHE[key] = entry
        entry.source_maps.copy_(source)
        entry.target_maps.copy_(target)
        entry.graph.replay()
        return DartelSolveResult(flow=entry.flow[0].clone(),
                                 source_maps=entry.source_maps.clone(),
                                 target_maps=entry.target_maps.clone(),
                                 metrics=entry.metrics.to(dtype).clone())
    flow = torch.empty((2, ny, nx), dtype=dtype, device=dev)
    metrics = torch.empty((steps, loop * its, 3), dtype=torch.float64, device=dev)
    ws = solve_workspace(nx, ny, steps, dtype, dev)
    _launch_solve(source, target, steps_per_job=steps, loop=loop, lmreg=lmreg, cycles=cycles,
                  nit=nit, its=its, code=code, assume_initial_nonzero=assume_initial_nonzero,
                  flow=flow, metrics=metrics, workspace=ws)
    return DartelSolveResult(flow=flow, source_maps=source, target_maps=target,
                             metrics=metrics.to(dtype))


def solve_dartel_maps_jobs(source_maps, target_maps, steps_per_job: int, *, loop: int = 6,
                           lmreg: float = 1e-3, cycles: int = 3, nit: int = 3, its: int = 3,
                           code: int = 1, device: str | torch.device = "auto",
                           dtype: torch.dtype = torch.float64,
                           assume_initial_nonzero: bool = False,
                           cuda_graph: bool = False) -> list[DartelSolveResult]:
    """Several independent solves in one batched launch.

    ``source_maps`` / ``target_maps`` are [jobs · steps_per_job, ny, nx]; each run
    of ``steps_per_job`` consecutive maps is one solve_dartel_maps job.  Every
    kernel of the solve processes all jobs at once; the results equal solving
    each job alone.  With ``cuda_graph`` the whole batch is captured once per
    shape and replayed.
    """

    dev = resolve_device(device)
    source = torch.as_tensor(source_maps, dtype=dtype, device=dev).contiguous()
    target = torch.as_tensor(target_maps, dtype=dtype, device=dev).contiguous()
    if source.ndim != 3 or tuple(source.shape) != tuple(target.shape):
        raise ValueError("source_maps and target_maps must have the same [batch, ny, nx] shape")
    if steps_per_job < 1 or source.shape[0] % steps_per_job:
        raise ValueError("the batch must be a whole number of jobs of steps_per_job maps")
    if cycles < 0 or nit < 0 or its < 0 or loop <= 0:
        raise ValueError("loop must be positive; cycles, nit and its non-negative")
    batch, ny, nx = (int(d) for d in source.shape)
    jobs = batch // steps_per_job
    if cuda_graph:
        # one graph (and one set of static buffers) per stream: solves of one
        # shape that run concurrently on different streams must not share them
        key = _solve_key(source, target, loop=loop, lmreg=float(lmreg), cycles=cycles, nit=nit,
                         its=its, code=code, steps_per_job=int(steps_per_job),
                         stream=int(torch.cuda.current_stream(dev).cuda_stream))
        entry = _DARTEL_CUDA_GRAPH_CACHE.get(key)
        if entry is None:
            entry = _capture(source, target, loop=loop, lmreg=lmreg, cycles=cycles, nit=nit,
                             its=its, code=code, device=dev, steps_per_job=steps_per_job)
            _DARTEL_CUDA_GRAPH_CACHE[key] = entry
        entry.source_maps.copy_(source)
        entry.target_maps.copy_(target)
        entry.graph.replay()
        flow, metrics = entry.flow.clone(), entry.metrics.clone()
    else:
        flow = torch.empty((jobs, 2, ny, nx), dtype=dtype, device=dev)
        metrics = torch.empty((batch, loop * its, 3), dtype=torch.float64, device=dev)
        ws = solve_workspace(nx, ny, batch, dtype, dev)
        _launch_solve(source, target, steps_per_job=steps_per_job, loop=loop, lmreg=lmreg,
                      cycles=cycles, nit=nit, its=its, code=code,
                      assume_initial_nonzero=assume_initial_nonzero, flow=flow,
                      metrics=metrics, workspace=ws)
    out = []
    for j in range(jobs):
        sl = slice(j * steps_per_job, (j + 1) * steps_per_job)
        out.append(DartelSolveResult(flow=flow[j], source_maps=source[sl],
                                     target_maps=target[sl], metrics=metrics[sl].to(dtype)))
    return out


def curvature_sheet_maps(vertices, stencil, fwhm: Sequence[float], curvtypes: Sequence[int],
                         dtype: torch.dtype, device: torch.device) -> torch.Tensor:
    """[steps, ny, nx] curvature sheets of one surface, build_maps of
    surface_warp.py:427-448 (resampling, then the per-step rule)."""

    from .surface_stencil import SurfaceStencil

    st = stencil.to(device) if isinstance(stencil, SurfaceStencil) else stencil
    v = torch.as_tensor(vertices, device=device).to(torch.float32).contiguous()
    return st.curvature_sheets(st.resample_vertices(v), tuple(float(f) for f in fwhm),
                               curvtypes, out_dtype=dtype)


def _check_curvtypes(curvtypes: Sequence[int], n_steps: int) -> tuple[int, ...]:
    if n_steps > len(curvtypes):
        raise ValueError("n_steps must not exceed the length of curvtypes")
    types = tuple(int(t) for t in curvtypes[:n_steps])
    for t in types:
        if t not in (2, 5):
            raise NotImplementedError("only upstream curvtype 2 and 5 exist")
    return types


_SIDE_STREAMS: dict[tuple, torch.cuda.Stream] = {}

# Concurrent batched solves per map shape in solve_dartel_from_surfaces_jobs
# (jobs split evenly between them; the results do not depend on the split).
_SOLVE_STREAMS = 1


def set_solve_streams(count: int) -> int:
    """How many concurrent solves the jobs of one map shape are split over;
    returns the previous setting."""

    global _SOLVE_STREAMS
    if int(count) < 1:
        raise ValueError("at least one solve stream is required")
    previous = _SOLVE_STREAMS
    _SOLVE_STREAMS = int(count)
    return previous


def _side_stream(device: torch.device, index: int) -> torch.cuda.Stream:
    """A reusable default-priority stream per (device, index)."""

    key = (device.type, device.index, index)
    if key not in _SIDE_STREAMS:
        _SIDE_STREAMS[key] = torch.cuda.Stream(device=device)
    return _SIDE_STREAMS[key]


def solve_dartel_from_surfaces_jobs(requests, *, n_steps: int = 2, loop: int = 6,
                                    lmreg: float = 1e-3, cycles: int = 3, nit: int = 3,
                                    its: int = 3, code: int = 1,
                                    curvtypes: Sequence[int] = (5, 5, 2),
                                    device: str | torch.device = "auto",
                                    dtype: torch.dtype = torch.float64,
                                    assume_initial_nonzero: bool = False,
                                    cuda_graph: bool = False,
                                    prepared_sides: dict | None = None
                                    ) -> list[DartelSolveResult]:
    """solve_dartel_from_surfaces for several independent registrations at once.

    ``requests`` holds (source_vertices, target_vertices, source_stencil,
    target_stencil, fwhm) tuples; their curvature sheets are computed per side
    (sides of several all-type-5 steps together in shared launches, any other
    step list side by side and step by step, surface_warp.py:427-448) and all
    solves run as one batched solve (solve_dartel_maps_jobs).

    ``prepared_sides`` (a dict the caller keeps across calls) holds the
    fwhm-independent type-5 preparation (resampled geometry, its fwhm-50 heat
    parameters and smoothed geometry) of each side, keyed by the identity of
    its vertex tensor and stencil: a side passed again with the same two
    objects (a registration's target in every DARTEL run) reuses it.  The
    preparation is a pure function of those two inputs, so the sheets are the
    same bits.
    """

    from .surface_stencil import SurfaceStencil, prepare_type5, type5_sheets_many

    types = _check_curvtypes(curvtypes, n_steps)
    dev = resolve_device(device)
    stencils, geoms, fws, origins = [], [], [], []
    for sv, tv, ss, ts, fw in requests:
        if len(fw) < n_steps:
            raise ValueError("each request needs at least n_steps fwhm values")
        for st, vert in ((ss, sv), (ts, tv)):
            origins.append((st, vert))
            st = st.to(dev) if isinstance(st, SurfaceStencil) else st
            stencils.append(st)
            geoms.append(None)
            fws.append(tuple(float(f) for f in fw[:n_steps]))
    all_type5 = n_steps > 1 and all(t == 5 for t in types)
    kept: list[tuple | None] = [None] * len(stencils)
    if all_type5 and prepared_sides is not None:
        for i, (st, vert) in enumerate(origins):
            entry = prepared_sides.get((id(st), id(vert)))
            if entry is not None and entry[0] is st and entry[1] is vert:
                kept[i] = entry[2:]
    for i, (st, vert) in enumerate(origins):
        if kept[i] is None:
            v = torch.as_tensor(vert, device=dev).to(torch.float32).contiguous()
            geoms[i] = stencils[i].resample_vertices(v)
    sheets: list[torch.Tensor | None] = [None] * len(stencils)
    if all_type5:
        # all sides that share the fwhm list are computed together
        by_fwhm: dict[tuple, list[int]] = {}
        for i, f in enumerate(fws):
            by_fwhm.setdefault(f, []).append(i)
        # the fwhm-50 geometry smoothing does not depend on the fwhms: every side's
        # in the same launches (sides prepared by an earlier call are reused)
        g32: list = [k[0] if k is not None else None for k in kept]
        heats: list = [k[1] if k is not None else None for k in kept]
        smoothed: list = [k[2] if k is not None else None for k in kept]
        todo = [i for i in range(len(stencils)) if kept[i] is None]
        if todo:
            pg, ph, ps = prepare_type5([stencils[i] for i in todo], [geoms[i] for i in todo])
            for k, i in enumerate(todo):
                g32[i], heats[i], smoothed[i] = pg[k], ph[k], ps[k]
                if prepared_sides is not None:
                    st, vert = origins[i]
                    prepared_sides[(id(st), id(vert))] = (st, vert, pg[k], ph[k], ps[k])
        for f, members in by_fwhm.items():
            res = type5_sheets_many([stencils[i] for i in members], [g32[i] for i in members],
                                    f, dtype, prepared=([g32[i] for i in members],
                                                        [heats[i] for i in members],
                                                        [smoothed[i] for i in members]))
            for i, r in zip(members, res):
                sheets[i] = r
    else:
        for i, (st, g, f) in enumerate(zip(stencils, geoms, fws)):
            sheets[i] = st.curvature_sheets(g, f, types, out_dtype=dtype)
    # one batched solve per map shape; shapes run on streams of their own so
    # that small maps overlap large ones
    shapes: dict[tuple, list[int]] = {}
    for k in range(len(requests)):
        shapes.setdefault(tuple(sheets[2 * k].shape[1:]), []).append(k)
    # the jobs of one map shape can be split over _SOLVE_STREAMS concurrent solves
    # (default 1: on the 61,442-vertex pair two streams measured no faster)
    groups: dict[tuple, list[int]] = {}
    for shape, members in shapes.items():
        parts = max(1, min(_SOLVE_STREAMS, len(members)))
        for p in range(parts):
            chunk = members[p * len(members) // parts:(p + 1) * len(members) // parts]
            if chunk:
                groups[shape + (p,)] = chunk
    results: list[DartelSolveResult | None] = [None] * len(requests)
    kw = dict(loop=loop, lmreg=lmreg, cycles=cycles, nit=nit, its=its, code=code, device=dev,
              dtype=dtype, assume_initial_nonzero=assume_initial_nonzero, cuda_graph=cuda_graph)
    if len(groups) == 1:
        members = next(iter(groups.values()))
        src = torch.cat([sheets[2 * k] for k in members])
        tgt = torch.cat([sheets[2 * k + 1] for k in members])
        return solve_dartel_maps_jobs(src, tgt, n_steps, **kw)
    current = torch.cuda.current_stream(dev)
    streams = []
    # (while a schedule is captured, the side streams' work joins the graph as
    # parallel branches: they wait on the capturing stream and it waits on them)
    for members in groups.values():
        src = torch.cat([sheets[2 * k] for k in members])
        tgt = torch.cat([sheets[2 * k + 1] for k in members])
        side = _side_stream(dev, len(streams))
        side.wait_stream(current)
        with torch.cuda.stream(side):
            out = solve_dartel_maps_jobs(src, tgt, n_steps, **kw)
        for k, r in zip(members, out):
            # the results are read on the caller's stream: keep their memory
            # from being reused on the side stream before that
            for tensor in (r.flow, r.source_maps, r.target_maps, r.metrics):
                tensor.record_stream(current)
            results[k] = r
        src.record_stream(side)
        tgt.record_stream(side)
        streams.append(side)
    for side in streams:
        current.wait_stream(side)
    return results


def solve_dartel_from_surfaces(source_vertices, target_vertices, source_stencil, target_stencil, *,
                               fwhm: Sequence[float] = (5.0, 5.0 / 3.0, 5.0 / 9.0),
                               curvtypes: Sequence[int] = (5, 5, 2), loop: int = 6,
                               n_steps: int = 2, lmreg: float = 1e-3, cycles: int = 3,
                               nit: int = 3, its: int = 3, code: int = 1, kernel: str = "auto",
                               squaring_kernel: str | None = None,
                               device: str | torch.device = "auto",
                               dtype: torch.dtype = torch.float64, parallel_sides: bool = True,
                               optimized: bool = True, assume_initial_nonzero: bool = False,
                               cuda_graph: bool = False) -> DartelSolveResult:
    """Curvature sheets of both surfaces, then their DARTEL flow (surface_warp.py:370-488).

    Sheets are produced directly in the solve dtype by the fused curvature
    kernel chain of each side.
    """

    from .surface_stencil import SurfaceStencil

    if n_steps < 1 or n_steps > len(fwhm) or n_steps > len(curvtypes):
        raise ValueError("n_steps must not exceed the lengths of fwhm and curvtypes")
    types = _check_curvtypes(curvtypes, n_steps)
    if squaring_kernel not in {None, "auto", "hip", "torch", "triton"}:
        raise ValueError("squaring_kernel must be None, 'auto', or 'hip'")
    dev = resolve_device(device)
    src_st = (source_stencil.to(dev) if isinstance(source_stencil, SurfaceStencil)
              else source_stencil)
    tgt_st = (target_stencil.to(dev) if isinstance(target_stencil, SurfaceStencil)
              else target_stencil)
    fw = tuple(float(f) for f in fwhm[:n_steps])
    src_maps = curvature_sheet_maps(source_vertices, src_st, fw, types, dtype, dev)
    tgt_maps = curvature_sheet_maps(target_vertices, tgt_st, fw, types, dtype, dev)
    return solve_dartel_maps(src_maps, tgt_maps, loop=loop, lmreg=lmreg, cycles=cycles, nit=nit,
                             its=its, code=code, kernel=kernel, squaring_kernel=squaring_kernel,
                             device=dev, dtype=dtype, optimized=optimized,
                             assume_initial_nonzero=assume_initial_nonzero, cuda_graph=cuda_graph)


def apply_flow_to_sphere(sphere_vertices, flow, *, unit_sphere_vertices=None,
                         inverse: bool = True, device: str | torch.device = "auto",
                         dtype: torch.dtype = torch.float64) -> torch.Tensor:
    """Warp sphere points through a flow (surface_warp.py:500-584)."""

    dev = resolve_device(device)
    sphere = torch.as_tensor(sphere_vertices, device=dev)
    deformation = torch.as_tensor(flow, dtype=dtype, device=dev).contiguous()
    if sphere.ndim != 2 or sphere.shape[1] != 3:
        raise ValueError("sphere_vertices must have shape [points, 3]")
    if deformation.ndim != 3 or deformation.shape[0] != 2:
        raise ValueError("flow must have shape [2, ny, nx]")
    if unit_sphere_vertices is None:
        unit = _normalise_f32(sphere.to(torch.float32).contiguous())
    else:
        unit = torch.as_tensor(unit_sphere_vertices, dtype=torch.float32, device=dev).contiguous()
        if unit.shape != sphere.shape:
            raise ValueError("unit_sphere_vertices must match sphere_vertices shape")
    n = int(unit.shape[0])
    out = torch.empty((n, 3), dtype=dtype, device=dev)
    ny, nx = int(deformation.shape[1]), int(deformation.shape[2])
    _lib.check(_lib.lib().csg_apply_flow(_lib.ptr(unit), n, _lib.ptr(deformation), nx, ny,
                                         _lib.dtype_code(dtype), int(bool(inverse)), _lib.ptr(out),
                                         _lib.stream_handle(dev)), "csg_apply_flow")
    return out


def _normalise_f32(points: torch.Tensor) -> torch.Tensor:
    """Per-vertex float32 unit vectors, used only when no stencil unit points are
    cached (surface_warp.py:518-522); the pipeline always passes cached points."""

    out = torch.empty_like(points)
    _lib.check(_lib.lib().csg_normalise_points(_lib.ptr(points), int(points.shape[0]),
                                               _lib.ptr(out), _lib.stream_handle(points.device)),
               "csg_normalise_points")
    return out


def apply_flow_to_stenciled_sphere(sphere_vertices, flow, stencil, *, inverse: bool = True,
                                   device: str | torch.device = "auto",
                                   dtype: torch.dtype = torch.float64) -> torch.Tensor:
    """Warp with the stencil's cached unit-sphere points (surface_warp.py:587-617)."""

    from .surface_stencil import SurfaceStencil

    dev = resolve_device(device)
    st = stencil.to(dev) if isinstance(stencil, SurfaceStencil) else stencil
    if st.unit_sphere_points is None:
        raise ValueError("stencil does not contain unit-sphere points required by apply_warp")
    return apply_flow_to_sphere(sphere_vertices, flow, unit_sphere_vertices=st.unit_sphere_points,
                                inverse=inverse, device=dev, dtype=dtype)
