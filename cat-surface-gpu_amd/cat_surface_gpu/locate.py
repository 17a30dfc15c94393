"""Device triangle index and point location (libcsg_hip).

Shared by the rotation search (surface_rotation.py) and the in-process
stencil builder (stencil_builder.py).  The index is the reference's uniform
grid over [-1, 1]^3 (src/cat_surface_gpu/surface_rotation.py:57-120): each
triangle is listed, in face order, in every cell of its bounding box grown by
``margin`` cells.  It is built on the GPU (per-face boxes, scan, emitted
(cell, face) pairs, stable radix sort by cell).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from . import _lib


@dataclass
class GridIndex:
    grid_size: int
    margin: int
    cell_off: torch.Tensor    # int32 [grid^3 + 1]
    cell_faces: torch.Tensor  # int32 [pairs]

    @property
    def max_candidates(self) -> int:
        return int(torch.diff(self.cell_off).max().item())


def build_grid_index(points: torch.Tensor, faces: torch.Tensor, grid_size: int,
                     margin: int) -> GridIndex:
    """Build the cell → triangle CSR on the device (synchronises once)."""

    if grid_size < 1 or margin < 0:
        raise ValueError("grid_size must be positive and margin non-negative")
    dev = points.device
    pts = points.contiguous()
    if pts.dtype not in (torch.float32, torch.float64):
        pts = pts.to(torch.float64)
    f = faces.to(torch.int32).contiguous()
    nf = int(f.shape[0])
    lib = _lib.lib()
    stream = _lib.stream_handle(dev)
    plan = _lib.workspace(lib.csg_grid_index_plan_workspace_bytes(nf), dev)
    npairs = ctypes.c_int()
    _lib.check(lib.csg_grid_index_plan(_lib.ptr(pts), _lib.dtype_code(pts.dtype), _lib.ptr(f), nf,
                                       int(grid_size), int(margin), _lib.ptr(plan), plan.numel(),
                                       ctypes.byref(npairs), stream), "csg_grid_index_plan")
    n = npairs.value
    if n < 1:
        raise ValueError("At least one triangle does not cover any spatial grid cell")
    cell_off = torch.empty(grid_size ** 3 + 1, dtype=torch.int32, device=dev)
    cell_faces = torch.empty(n, dtype=torch.int32, device=dev)
    ws = _lib.workspace(lib.csg_grid_index_fill_workspace_bytes(n), dev)
    _lib.check(lib.csg_grid_index_fill(_lib.ptr(plan), nf, int(grid_size), n, _lib.ptr(cell_off),
                                       _lib.ptr(cell_faces), _lib.ptr(ws), ws.numel(), stream),
               "csg_grid_index_fill")
    return GridIndex(int(grid_size), int(margin), cell_off, cell_faces)


def closest_triangles(queries: torch.Tensor, verts: torch.Tensor, faces: torch.Tensor, *,
                      grid_size: int, margin: int = 1, max_margin: int = 8):
    """Exact closest triangle and barycentric weights for each query (float64).

    The search covers the query's cell; a query farther than ``margin`` cells
    from its best triangle could have its true closest triangle elsewhere, so
    the index is rebuilt with a larger margin until no query is that far.
    Returns (face [n] int32, vertex indices [n, 3] int32, weights [n, 3] f64).
    """

    q = queries.to(torch.float64)
    v = verts.to(torch.float64)
    # the grid spans [-1, 1]^3: scale both point sets into it when they reach
    # outside (argmin and weights are invariant under a common scale); decided
    # on the device, without a host round trip
    extent = torch.maximum(q.abs().max(), v.abs().max())
    scale = torch.where(extent > 1.0, extent * 1.0001, torch.ones_like(extent))
    q = (q / scale).contiguous()
    v = (v / scale).contiguous()
    f = faces.to(torch.int32).contiguous()
    nq = int(q.shape[0])
    dev = q.device
    lib = _lib.lib()
    face = torch.empty(nq, dtype=torch.int32, device=dev)
    idx = torch.empty((nq, 3), dtype=torch.int32, device=dev)
    w = torch.empty((nq, 3), dtype=torch.float64, device=dev)
    ws = _lib.workspace(256, dev)
    while True:
        index = build_grid_index(v, f, grid_size, margin)
        far = ctypes.c_int()
        reach = margin * 2.0 / grid_size
        _lib.check(lib.csg_closest_triangles(_lib.ptr(q), nq, _lib.ptr(v), _lib.ptr(f),
                                             _lib.ptr(index.cell_off), _lib.ptr(index.cell_faces),
                                             grid_size, reach, _lib.ptr(face), _lib.ptr(idx),
                                             _lib.ptr(w), ctypes.byref(far), _lib.ptr(ws),
                                             ws.numel(), _lib.stream_handle(dev)),
                   "csg_closest_triangles")
        if far.value == 0:
            return face, idx, w
        if margin >= max_margin:
            raise RuntimeError(f"{far.value} points lie farther than {reach} from the mesh")
        margin += 1
