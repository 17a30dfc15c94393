"""High-level initial-rotation pipeline (src/cat_surface_gpu/rotation_pipeline.py)."""

from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path

import numpy as np
import torch

from .dartel_grid import resolve_device
from .surface_rotation import (
    RotationCostInputs,
    RotationGridIndex,
    RotationGridIndexDevice,
    RotationSearchResult,
)
from .surface_stencil import SurfaceStencil


def read_rotation_values(path: str | Path) -> tuple[np.ndarray, np.ndarray]:
    """[i32 ns, i32 nt, f64 ns, f64 nt] (rotation_pipeline.py:23-38)."""

    path = Path(path)
    with path.open("rb") as stream:
        header = np.fromfile(stream, dtype="<i4", count=2)
        if header.size != 2:
            raise ValueError(f"Rotation feature file header is short: {path}")
        source = np.fromfile(stream, dtype="<f8", count=int(header[0]))
        target = np.fromfile(stream, dtype="<f8", count=int(header[1]))
        trailing = stream.read(1)
    if source.size != int(header[0]) or target.size != int(header[1]):
        raise ValueError(f"Rotation feature file length is invalid: {path}")
    if trailing:
        raise ValueError(f"Rotation feature file has trailing data: {path}")
    return np.ascontiguousarray(source), np.ascontiguousarray(target)


def write_rotation_values(path: str | Path, source: np.ndarray, target: np.ndarray) -> None:
    with Path(path).open("wb") as stream:
        np.asarray([source.size, target.size], dtype="<i4").tofile(stream)
        np.asarray(source, dtype="<f8").tofile(stream)
        np.asarray(target, dtype="<f8").tofile(stream)


def read_rotation_points(path: str | Path) -> np.ndarray:
    """[i32 n] then n×3 float32 or float64 (rotation_pipeline.py:41-66)."""

    path = Path(path)
    with path.open("rb") as stream:
        header = np.fromfile(stream, dtype="<i4", count=1)
        if header.size != 1 or int(header[0]) < 0:
            raise ValueError(f"Rotation points file header is invalid: {path}")
        n = int(header[0])
        payload = stream.read()
    if len(payload) == 12 * n:
        pts = np.frombuffer(payload, dtype="<f4", count=3 * n)
    elif len(payload) == 24 * n:
        pts = np.frombuffer(payload, dtype="<f8", count=3 * n)
    else:
        raise ValueError(f"Rotation points file length is invalid: {path}, {len(payload)}")
    return np.ascontiguousarray(pts.reshape(n, 3))


@dataclass
class RotationPipeline:
    """Reusable target index plus the CAT-compatible search."""

    index_cpu: RotationGridIndex | None
    index: RotationGridIndexDevice

    @classmethod
    def from_stencil(cls, target_stencil, *, device: str | torch.device = "auto",
                     grid_size: int = 128, margin: int = 1,
                     candidate_table_dtype: torch.dtype = torch.int32,
                     geometry_dtype: torch.dtype = torch.float32,
                     compressed_candidate_table: bool = True) -> "RotationPipeline":
        """Build the index on the device from a host or device stencil."""

        dev = resolve_device(device)
        if isinstance(target_stencil, SurfaceStencil):
            pts = torch.as_tensor(target_stencil.sphere_points, dtype=torch.float64, device=dev)
            faces = torch.as_tensor(target_stencil.faces, dtype=torch.int32, device=dev)
        else:
            pts = target_stencil.sphere_points64
            if pts is None:
                pts = target_stencil.sphere_points.to(torch.float64)
            faces = target_stencil.faces
        index = RotationGridIndexDevice.from_device_geometry(pts, faces, grid_size=grid_size,
                                                             margin=margin)
        return cls(index_cpu=None, index=index)

    @property
    def device(self) -> torch.device:
        return self.index.device

    def prepare_inputs(self, source_points, source_values, target_values) -> RotationCostInputs:
        return self.index.prepare_cost_inputs(source_points, source_values, target_values)

    def search(self, source_points, source_values, target_values, *, point_chunk: int = 1024,
               refine: bool = True, max_iter: int = 500, tol: float = 1.0e-4,
               simplex_step: float = 0.1) -> RotationSearchResult:
        inputs = self.prepare_inputs(source_points, source_values, target_values)
        return self.index.search_rotation_prepared(inputs, point_chunk=point_chunk, refine=refine,
                                                   max_iter=max_iter, tol=tol,
                                                   simplex_step=simplex_step)


def run_rotation_pipeline(target_stencil, source_points, source_values, target_values, *,
                          device: str | torch.device = "auto", grid_size: int = 128,
                          margin: int = 1, candidate_table_dtype: torch.dtype = torch.int32,
                          geometry_dtype: torch.dtype = torch.float32, point_chunk: int = 1024,
                          refine: bool = True, max_iter: int = 500, tol: float = 1.0e-4,
                          simplex_step: float = 0.1) -> RotationSearchResult:
    pipeline = RotationPipeline.from_stencil(target_stencil, device=device, grid_size=grid_size,
                                             margin=margin)
    return pipeline.search(source_points, source_values, target_values, point_chunk=point_chunk,
                           refine=refine, max_iter=max_iter, tol=tol, simplex_step=simplex_step)


def run_rotation_pipeline_from_files(source_points_path, target_stencil_path,
                                     rotation_values_path, **kwargs) -> RotationSearchResult:
    source_points = read_rotation_points(source_points_path)
    source_values, target_values = read_rotation_values(rotation_values_path)
    target_stencil = SurfaceStencil.from_file(target_stencil_path)
    return run_rotation_pipeline(target_stencil, source_points, source_values, target_values,
                                 **kwargs)
