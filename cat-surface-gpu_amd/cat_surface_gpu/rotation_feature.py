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

from . import _lib
from .surface_stencil import SurfaceStencilDevice, _minmax, heat_parameters_from_edge_mean


@dataclass(frozen=True)
class RotationFeatureResult:
    values: torch.Tensor
    iterations: int
    relative_residual: float


def depth_potential(stencil: SurfaceStencilDevice, smoothed_surface: torch.Tensor, *,
                    alpha: float = 1.0 / 1000.0, max_iter: int = 1000,
                    relaxation: float = 1.90) -> tuple[torch.Tensor, int, float]:
    """Depth potential by colour-ordered SOR (rotation_feature.py:253-317)."""

    topo = stencil.topology
    pts = smoothed_surface.to(torch.float32).contiguous()
    n = int(pts.shape[0])
    nf = int(topo.faces.shape[0])
    nnz = int(topo.sys_cols.numel())
    x = torch.empty(n, dtype=torch.float64, device=pts.device)
    rel = torch.empty(1, dtype=torch.float64, device=pts.device)
    lib = _lib.lib()
    ws = _lib.workspace(lib.csg_depth_workspace_bytes(n, nf, nnz), pts.device)
    coff = (ctypes.c_int * topo.colour_off.size)(*topo.colour_off.tolist())
    _lib.check(lib.csg_depth_potential(
        _lib.ptr(pts), n, _lib.ptr(topo.faces), nf, _lib.ptr(topo.inc_off), _lib.ptr(topo.inc),
        _lib.ptr(topo.sys_row_off), _lib.ptr(topo.sys_cols), _lib.ptr(topo.sys_perm),
        _lib.ptr(topo.sys_alpha), nnz, coff, _lib.ptr(topo.colour_verts),
        int(topo.colour_off.size - 1), float(alpha), int(max_iter), float(relaxation),
        _lib.ptr(x), _lib.ptr(rel), _lib.ptr(ws), ws.numel(), _lib.stream_handle(pts.device)),
        "csg_depth_potential")
    return x, int(max_iter), float(rel.item())


def compute_rotation_feature(stencil: SurfaceStencilDevice, surface_vertices: torch.Tensor, *,
                             heat_fwhm: float, smoothed_surface: torch.Tensor | None = None,
                             depth_values: torch.Tensor | None = None,
                             depth_alpha: float = 1.0 / 1000.0, max_iter: int = 1000,
                             tolerance: float = 1.0e-9,
                             solver: str = "colored-sor") -> RotationFeatureResult:
    """One CAT-compatible smoothed depth-potential feature (rotation_feature.py:320-399)."""

    if solver not in {"colored-sor", "pcg"}:
        raise ValueError("rotation feature solver must be 'colored-sor' or 'pcg'")
    if smoothed_surface is None:
        mapped = stencil.resample_vertices(surface_vertices)
        smoothed_surface = stencil.smooth_geometry(mapped, heat_fwhm)
    else:
        smoothed_surface = torch.as_tensor(smoothed_surface, dtype=torch.float32,
                                           device=stencil.device).contiguous()
        expected = (int(stencil.sphere_points.shape[0]), 3)
        if tuple(smoothed_surface.shape) != expected:
            raise ValueError(f"coarse-feature geometry must have shape {expected}")
    if depth_values is not None:
        values = torch.as_tensor(depth_values, dtype=torch.float64,
                                 device=stencil.device).reshape(-1).contiguous()
        if values.numel() != stencil.sphere_points.shape[0]:
            raise ValueError("Upstream depth-potential length must match the coarse sphere")
        iterations, residual = 0, 0.0
    else:
        if solver == "pcg":
            raise NotImplementedError(
                "this build solves the depth potential with the reference's default "
                "colour-ordered SOR; solver='pcg' is not built")
        values, iterations, residual = depth_potential(stencil, smoothed_surface,
                                                       alpha=depth_alpha, max_iter=max_iter)
    em = stencil.edge_mean_distance(smoothed_surface)
    params = [heat_parameters_from_edge_mean(em, 50.0)]
    smoothed = stencil._smooth_values_batch(values.reshape(1, -1), smoothed_surface, params,
                                            False)[0]
    return RotationFeatureResult(values=_minmax(smoothed, 2.220446049250313e-16),
                                 iterations=iterations, relative_residual=residual)


def compute_rotation_features(stencils: Sequence[SurfaceStencilDevice],
                              surfaces: Sequence[torch.Tensor], *,
                              heat_fwhms: Sequence[float] = (15.0, 10.0),
                              smoothed_surfaces=None, depth_values=None,
                              depth_alpha: float = 1.0 / 1000.0, max_iter: int = 1000,
                              tolerance: float = 1.0e-9, solver: str = "colored-sor",
                              parallel_sides: bool = True) -> tuple[RotationFeatureResult, ...]:
    """Source and target features (rotation_feature.py:402-458), one after the other
    on the current stream (each side already fills the device)."""

    if len(stencils) != len(surfaces) or len(stencils) != len(heat_fwhms):
        raise ValueError("stencils, surfaces, and heat_fwhms must have equal lengths")
    if smoothed_surfaces is not None and len(smoothed_surfaces) != len(stencils):
        raise ValueError("smoothed_surfaces length must match stencils")
    if depth_values is not None and len(depth_values) != len(stencils):
        raise ValueError("depth_values length must match stencils")
    return tuple(
        compute_rotation_feature(
            stencils[i], surfaces[i], heat_fwhm=float(heat_fwhms[i]),
            smoothed_surface=None if smoothed_surfaces is None else smoothed_surfaces[i],
            depth_values=None if depth_values is None else depth_values[i],
            depth_alpha=depth_alpha, max_iter=max_iter, tolerance=tolerance, solver=solver)
        for i in range(len(stencils)))
