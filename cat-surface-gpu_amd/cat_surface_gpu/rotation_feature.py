"""Rotation features: smoothed depth potential on the stencil sphere (MI355X).

Drop-in for src/cat_surface_gpu/rotation_feature.py.  The "cuda" backend of
the reference (resample → heat-smooth geometry → depth potential by
colour-ordered SOR → fwhm-50 value smoothing → min/max) runs entirely in
libcsg_hip.so; given helper depth values (the reference's official-depth
backend) only the last two stages run.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Sequence

import torch

from . import _lib, schedule
from .surface_stencil import (SurfaceStencilDevice, _minmax, edge_heat,
                              heat_parameters_from_edge_mean, smooth_geometry_many, sor_patches)


@dataclass(frozen=True)
class RotationFeatureResult:
    values: torch.Tensor
    iterations: int
    relative_residual: float


def set_sor_placement(mode) -> int:
    """Choose where the depth-potential SOR sweeps run; returns the previous mode.

    ``2`` (the default) keeps each mesh's rows on the workgroups of one XCD so
    that colours synchronise through that XCD's L2; ``1`` does the same with x
    stores written through to memory; ``0`` uses one device-wide grid barrier
    per colour.  All compute identical values.
    """

    code = int(mode)
    if code not in (0, 1, 2):
        raise ValueError("SOR placement must be 0, 1 or 2")
    return int(_lib.lib().csg_set_sor_placement(code))


# Depth-potential sweeps blocked in time on mesh patches (csg_depth_potential_patched);
# False: one synchronisation per colour (csg_depth_potential).  Identical values.
_SOR_PATCHES = True


def set_sor_patches(enabled: bool) -> bool:
    """Select the patch-blocked SOR schedule (default) or the per-colour one;
    returns the previous setting."""

    global _SOR_PATCHES
    previous = _SOR_PATCHES
    _SOR_PATCHES = bool(enabled)
    return previous


def depth_potential(stencil: SurfaceStencilDevice, smoothed_surface, *,
                    alpha: float = 1.0 / 1000.0, max_iter: int = 1000,
                    relaxation: float = 1.90, solver: str = "colored-sor",
                    tolerance: float = 1.0e-9) -> tuple[torch.Tensor, int, torch.Tensor]:
    """Depth potential by colour-ordered SOR (rotation_feature.py:253-317) or, with
    ``solver="pcg"``, Jacobi-preconditioned CG to ``tolerance`` (:193-250).

    ``smoothed_surface`` is one [n, 3] geometry or a [batch, n, 3] stack of
    geometries on the stencil's topology; all are solved in one persistent
    launch.  Returns the values ([n] or [batch, n], float64), the iteration
    count (SOR: max_iter; PCG: the largest of the batch) and the relative
    residuals |rhs - A x| / |rhs| (float64, on the device, shape [] or [batch]).
    """

    if solver not in {"colored-sor", "pcg"}:
        raise ValueError("rotation feature solver must be 'colored-sor' or 'pcg'")
    if solver == "pcg":
        return _depth_potential_pcg(stencil, smoothed_surface, alpha=alpha, max_iter=max_iter,
                                    tolerance=tolerance)

    topo = stencil.topology
    pts = torch.as_tensor(smoothed_surface, dtype=torch.float32,
                          device=stencil.device).contiguous()
    single = pts.dim() == 2
    if single:
        pts = pts.unsqueeze(0)
    if pts.dim() != 3 or pts.shape[2] != 3 or pts.shape[1] != topo.n_points:
        raise ValueError(f"depth-potential geometry must have shape [batch, {topo.n_points}, 3]")
    batch, n = int(pts.shape[0]), int(pts.shape[1])
    nf = int(topo.faces.shape[0])
    nnz = int(topo.sys_cols.numel())
    x = torch.empty(batch, n, dtype=torch.float64, device=pts.device)
    rel = torch.empty(batch, dtype=torch.float64, device=pts.device)
    lib = _lib.lib()
    ws = _lib.workspace(lib.csg_depth_workspace_bytes(n, nf, nnz, batch, topo.ell_width),
                        pts.device)
    coff = (ctypes.c_int * topo.colour_off.size)(*topo.colour_off.tolist())
    patches = None
    if _SOR_PATCHES and topo.colour_order_np is not None:
        # one patch per CU of an XCD: all of a mesh's patches run on its XCD at once
        per_xcd = max(1, torch.cuda.get_device_properties(pts.device).multi_processor_count // 8)
        # the partition follows the stencil's sphere (compact patches)
        patches = sor_patches(topo, stencil.sphere_points, core=-(-n // per_xcd)).abi()
    args = (_lib.ptr(pts), batch, n, _lib.ptr(topo.faces), nf, _lib.ptr(topo.inc_off),
            _lib.ptr(topo.inc), _lib.ptr(topo.sys_row_off), _lib.ptr(topo.sys_cols),
            _lib.ptr(topo.sys_contrib_off), _lib.ptr(topo.sys_contrib), nnz, coff,
            _lib.ptr(topo.colour_verts), int(topo.colour_off.size - 1), _lib.ptr(topo.ell_cols),
            _lib.ptr(topo.ell_src), int(topo.ell_width), float(alpha), int(max_iter),
            float(relaxation), None if patches is None else ctypes.byref(patches), _lib.ptr(x),
            _lib.ptr(rel))
    if schedule.capturing():
        # nothing may wait for the host: the time-out and XCD-placement verdicts
        # are checked after the graph's replay
        status = torch.zeros(2, dtype=torch.int32, device=pts.device)
        _lib.check(lib.csg_depth_potential_patched_deferred(
            *args, _lib.ptr(status), _lib.ptr(ws), ws.numel(), _lib.stream_handle(pts.device)),
            "csg_depth_potential_patched_deferred")
        schedule.defer(status, "depth-potential SOR timed out or was misplaced")
    else:
        _lib.check(lib.csg_depth_potential_patched(
            *args, _lib.ptr(ws), ws.numel(), _lib.stream_handle(pts.device)),
            "csg_depth_potential")
    if single:
        return x[0], int(max_iter), rel[0]
    return x, int(max_iter), rel


def _depth_potential_pcg(stencil: SurfaceStencilDevice, smoothed_surface, *, alpha: float,
                         max_iter: int, tolerance: float):
    topo = stencil.topology
    pts = torch.as_tensor(smoothed_surface, dtype=torch.float32,
                          device=stencil.device).contiguous()
    single = pts.dim() == 2
    if single:
        pts = pts.unsqueeze(0)
    if pts.dim() != 3 or pts.shape[2] != 3 or pts.shape[1] != topo.n_points:
        raise ValueError(f"depth-potential geometry must have shape [batch, {topo.n_points}, 3]")
    batch, n = int(pts.shape[0]), int(pts.shape[1])
    nf = int(topo.faces.shape[0])
    nnz = int(topo.sys_cols.numel())
    x = torch.empty(batch, n, dtype=torch.float64, device=pts.device)
    rel = torch.empty(batch, dtype=torch.float64, device=pts.device)
    iters = (ctypes.c_int * batch)()
    lib = _lib.lib()
    ws = _lib.workspace(lib.csg_depth_pcg_workspace_bytes(n, nf, nnz, batch), pts.device)
    _lib.check(lib.csg_depth_potential_pcg(
        _lib.ptr(pts), batch, n, _lib.ptr(topo.faces), nf, _lib.ptr(topo.inc_off),
        _lib.ptr(topo.inc), _lib.ptr(topo.sys_row_off), _lib.ptr(topo.sys_cols),
        _lib.ptr(topo.sys_contrib_off), _lib.ptr(topo.sys_contrib), nnz, float(alpha),
        int(max_iter), float(tolerance), _lib.ptr(x), _lib.ptr(rel), iters, _lib.ptr(ws),
        ws.numel(), _lib.stream_handle(pts.device)), "csg_depth_potential_pcg")
    its = [int(v) for v in iters]
    if single:
        return x[0], its[0], rel[0]
    return x, max(its), rel


def _feature_geometry(stencil: SurfaceStencilDevice, surface_vertices, heat_fwhm: float,
                      smoothed_surface) -> torch.Tensor:
    """Resampled, heat-smoothed geometry on the stencil sphere (rotation_feature.py:356-370)."""

    if smoothed_surface is None:
        mapped = stencil.resample_vertices(surface_vertices)
        return stencil.smooth_geometry(mapped, heat_fwhm)
    smoothed_surface = torch.as_tensor(smoothed_surface, dtype=torch.float32,
                                       device=stencil.device).contiguous()
    expected = (int(stencil.sphere_points.shape[0]), 3)
    if tuple(smoothed_surface.shape) != expected:
        raise ValueError(f"coarse-feature geometry must have shape {expected}")
    return smoothed_surface


def _given_depth(stencil: SurfaceStencilDevice, depth_values) -> torch.Tensor:
    values = torch.as_tensor(depth_values, dtype=torch.float64,
                             device=stencil.device).reshape(-1).contiguous()
    if values.numel() != stencil.sphere_points.shape[0]:
        raise ValueError("Upstream depth-potential length must match the coarse sphere")
    return values


def _finish_feature(stencil: SurfaceStencilDevice, geometry: torch.Tensor, values: torch.Tensor,
                    iterations: int, residual: float,
                    edge_mean: float | None = None) -> RotationFeatureResult:
    """fwhm-50 value smoothing and min/max normalisation (rotation_feature.py:384-399)."""

    em = stencil.edge_mean_distance(geometry) if edge_mean is None else edge_mean
    params = [heat_parameters_from_edge_mean(em, 50.0)]
    smoothed = stencil._smooth_values_batch(values.reshape(1, -1), geometry, params, False)[0]
    return RotationFeatureResult(values=_minmax(smoothed, 2.220446049250313e-16),
                                 iterations=iterations, relative_residual=residual)


def _finish_features(stencils: Sequence[SurfaceStencilDevice], geometries, values, iterations,
                     residuals, heats, groups) -> tuple[RotationFeatureResult, ...]:
    """_finish_feature for many sides: one smoothing and one normalisation per topology
    group, the fwhm-50 parameters (heats: EdgeHeat per side) read on the device."""

    out: list[RotationFeatureResult | None] = [None] * len(stencils)
    for g in groups:
        its = torch.cat([heats[i].iters for i in g])
        sig = torch.cat([heats[i].sigma for i in g])
        if len(g) > 1:
            geom = torch.stack([geometries[i] for i in g])
            vals = torch.stack([values[i].reshape(-1) for i in g])
            smoothed = stencils[g[0]]._smooth_values_batch_dev(vals, geom, its, sig, False,
                                                               mesh_of=range(len(g)))
        else:
            smoothed = stencils[g[0]]._smooth_values_batch_dev(values[g[0]].reshape(1, -1),
                                                               geometries[g[0]], its, sig,
                                                               False)
        normed = _minmax(smoothed, 2.220446049250313e-16)
        for k, i in enumerate(g):
            out[i] = RotationFeatureResult(values=normed[k], iterations=iterations[i],
                                           relative_residual=residuals[i])
    return tuple(out)


def _topology_groups(stencils: Sequence[SurfaceStencilDevice], together: bool) -> list[list[int]]:
    """Sides grouped by shared sphere topology (the cache returns one object per mesh)."""

    groups: list[list[int]] = []
    for i, st in enumerate(stencils):
        for g in groups:
            if together and len(g) < 32 and stencils[g[0]].topology is st.topology:
                g.append(i)
                break
        else:
            groups.append([i])
    return groups


def _check_solver(solver: str, need_solve: bool) -> None:
    if solver not in {"colored-sor", "pcg"}:
        raise ValueError("rotation feature solver must be 'colored-sor' or 'pcg'")


def compute_rotation_feature(stencil: SurfaceStencilDevice, surface_vertices: torch.Tensor, *,
                             heat_fwhm: float, smoothed_surface: torch.Tensor | None = None,
                             depth_values: torch.Tensor | None = None,
                             depth_alpha: float = 1.0 / 1000.0, max_iter: int = 1000,
                             tolerance: float = 1.0e-9,
                             solver: str = "colored-sor") -> RotationFeatureResult:
    """One CAT-compatible smoothed depth-potential feature (rotation_feature.py:320-399)."""

    _check_solver(solver, depth_values is None)
    geometry = _feature_geometry(stencil, surface_vertices, heat_fwhm, smoothed_surface)
    if depth_values is not None:
        return _finish_feature(stencil, geometry, _given_depth(stencil, depth_values), 0, 0.0)
    values, iterations, residual = depth_potential(stencil, geometry, alpha=depth_alpha,
                                                   max_iter=max_iter, solver=solver,
                                                   tolerance=tolerance)
    return _finish_feature(stencil, geometry, values, iterations, float(residual))


def compute_rotation_features(stencils: Sequence[SurfaceStencilDevice],
                              surfaces: Sequence[torch.Tensor], *,
                              heat_fwhms: Sequence[float] = (15.0, 10.0),
                              smoothed_surfaces=None, depth_values=None,
                              depth_alpha: float = 1.0 / 1000.0, max_iter: int = 1000,
                              tolerance: float = 1.0e-9, solver: str = "colored-sor",
                              parallel_sides: bool = True) -> tuple[RotationFeatureResult, ...]:
    """Source and target features (rotation_feature.py:402-458).

    Sides whose stencils share one sphere topology (the usual case: both are
    built on the same icosphere) have their depth potentials solved together
    in one persistent SOR launch; ``parallel_sides=False`` solves them one at a
    time.  The results are identical either way.
    """

    if len(stencils) != len(surfaces) or len(stencils) != len(heat_fwhms):
        raise ValueError("stencils, surfaces, and heat_fwhms must have equal lengths")
    if smoothed_surfaces is not None and len(smoothed_surfaces) != len(stencils):
        raise ValueError("smoothed_surfaces length must match stencils")
    if depth_values is not None and len(depth_values) != len(stencils):
        raise ValueError("depth_values length must match stencils")
    _check_solver(solver, depth_values is None)
    if smoothed_surfaces is None:
        # resample every side, then one host synchronisation for the edge means
        # and shared smoothing launches (the same values as per-side calls)
        mapped = [st.resample_vertices(torch.as_tensor(v, device=st.device))
                  for st, v in zip(stencils, surfaces)]
        heats = edge_heat(stencils, mapped, [(float(f),) for f in heat_fwhms])
        geometries = smooth_geometry_many(stencils, mapped, [h.iters[0:1] for h in heats])
    else:
        geometries = [_feature_geometry(stencils[i], surfaces[i], float(heat_fwhms[i]),
                                        smoothed_surfaces[i]) for i in range(len(stencils))]
    final_heats = edge_heat(stencils, geometries, [(50.0,)] * len(stencils))
    groups = _topology_groups(stencils, parallel_sides)
    if depth_values is not None:
        given = [_given_depth(stencils[i], depth_values[i]) for i in range(len(stencils))]
        zeros = [0] * len(stencils)
        return _finish_features(stencils, geometries, given, zeros, [0.0] * len(stencils),
                                final_heats, groups)
    values: list[torch.Tensor | None] = [None] * len(stencils)
    residuals: list[float] = [0.0] * len(stencils)
    iterations: list[int] = [int(max_iter)] * len(stencils)
    for g in groups:
        stack = torch.stack([geometries[i] for i in g]) if len(g) > 1 else geometries[g[0]]
        x, its, rel = depth_potential(stencils[g[0]], stack, alpha=depth_alpha,
                                      max_iter=max_iter, solver=solver, tolerance=tolerance)
        # a captured schedule cannot read the residuals (not used downstream)
        rel_host = ([float("nan")] * len(g) if schedule.capturing()
                    else rel.reshape(-1).tolist())
        for k, i in enumerate(g):
            values[i] = x[k] if len(g) > 1 else x
            residuals[i] = float(rel_host[k])
            iterations[i] = int(its)
    return _finish_features(stencils, geometries, values, iterations, residuals, final_heats,
                            groups)
