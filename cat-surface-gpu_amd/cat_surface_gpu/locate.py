"""Device triangle index and point location (libcsg_hip).

Shared by the rotation search (surface_rotation.py) and the in-process
stencil builder (stencil_builder.py).  The index is the reference's uniform
grid over [-1, 1]^3 (src/cat_surface_gpu/surface_rotation.py:57-120): each
triangle is listed, in face order, in every cell of its bounding box grown by
``margin`` cells.  It is built on the GPU: the exact two-phase build (per-face
boxes, scan, emitted (cell, face) pairs, stable radix sort by cell) and the
asynchronous one (per-cell counts, scan, fill, faces ranked within each cell).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from . import _lib

# Pair capacities for asynchronous index builds, learned from the exact counts
# of earlier builds: (faces, grid size, margin) -> capacity.
_CAPACITY: dict[tuple[int, int, int], int] = {}


def _learn_capacity(key: tuple[int, int, int], pairs: int) -> None:
    want = int(pairs * 1.08) + 1024
    if want > _CAPACITY.get(key, 0):
        _CAPACITY[key] = want


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
    _learn_capacity((nf, int(grid_size), int(margin)), n)
    cell_off = torch.empty(grid_size ** 3 + 1, dtype=torch.int32, device=dev)
    cell_faces = torch.empty(n, dtype=torch.int32, device=dev)
    ws = _lib.workspace(lib.csg_grid_index_fill_workspace_bytes(n), dev)
    _lib.check(lib.csg_grid_index_fill(_lib.ptr(plan), nf, int(grid_size), n, _lib.ptr(cell_off),
                                       _lib.ptr(cell_faces), _lib.ptr(ws), ws.numel(), stream),
               "csg_grid_index_fill")
    return GridIndex(int(grid_size), int(margin), cell_off, cell_faces)


def build_grid_index_async(points: torch.Tensor, faces: torch.Tensor, grid_size: int,
                           margin: int, capacity: int, required: torch.Tensor) -> GridIndex:
    """The index in one asynchronous call (csg_grid_index_build): at most
    ``capacity`` pairs; ``required`` (a device int) receives the count the index
    needs, and the index is complete iff that count is at most ``capacity``."""

    dev = points.device
    pts = points.contiguous()
    if pts.dtype not in (torch.float32, torch.float64):
        pts = pts.to(torch.float64)
    f = faces.to(torch.int32).contiguous()
    nf = int(f.shape[0])
    lib = _lib.lib()
    cell_off = torch.empty(grid_size ** 3 + 1, dtype=torch.int32, device=dev)
    cell_faces = torch.empty(capacity, dtype=torch.int32, device=dev)
    ws = _lib.workspace(lib.csg_grid_index_build_workspace_bytes(nf, capacity, int(grid_size)),
                        dev)
    _lib.check(lib.csg_grid_index_build(_lib.ptr(pts), _lib.dtype_code(pts.dtype), _lib.ptr(f), nf,
                                        int(grid_size), int(margin), int(capacity),
                                        _lib.ptr(cell_off), _lib.ptr(cell_faces),
                                        _lib.ptr(required), _lib.ptr(ws), ws.numel(),
                                        _lib.stream_handle(dev)), "csg_grid_index_build")
    return GridIndex(int(grid_size), int(margin), cell_off, cell_faces)


# Closest-triangle queries skip candidates by their faces' bounding spheres
# (csg_closest_triangles_pruned); False: every candidate is evaluated
# (csg_closest_triangles).  Identical results.
_PRUNE = True


def set_closest_pruning(enabled: bool) -> bool:
    """Select bounding-sphere pruning of the closest-triangle queries; returns the
    previous setting."""

    global _PRUNE
    previous = _PRUNE
    _PRUNE = bool(enabled)
    return previous


@dataclass
class LocateCheck:
    """What an asynchronous closest_triangles still has to confirm: status (device
    int32 [2]) = (pairs the index needed, queries beyond the searched reach)."""

    status: torch.Tensor
    capacity: int
    key: tuple[int, int, int]

    def passed(self, host_status) -> bool:
        required, far = int(host_status[0]), int(host_status[1])
        if required > self.capacity:
            _learn_capacity(self.key, required)
        return required <= self.capacity and far == 0


def closest_triangles(queries: torch.Tensor, verts: torch.Tensor, faces: torch.Tensor, *,
                      grid_size: int, margin: int = 1, max_margin: int = 8,
                      deferred: bool = False, block: bool = False):
    """Exact closest triangle and barycentric weights for each query (float64).

    The search covers the query's cell; a query farther than ``margin`` cells
    from its best triangle could have its true closest triangle elsewhere, so
    the index is rebuilt with a larger margin until no query is that far.
    With ``block=True`` (for queries on or near the mesh) the index lists each
    face only in the cells its bounding box overlaps and a query searches the
    block of cells within ``margin`` of its own only when its own cell cannot
    decide (csg_closest_triangles_block): far fewer index pairs, the same
    result; ``margin`` then grows the block instead of the index.
    Returns (face [n] int32, vertex indices [n, 3] int32, weights [n, 3] f64).

    ``deferred=True`` returns a fourth item: None when the result is final, else
    a LocateCheck; then nothing synchronises (the index is built with a pair
    capacity learned from earlier builds of the same face count, grid and
    margin, and only ``margin`` is searched) and the result is exact iff the
    check passes, which the caller confirms after its next synchronisation
    (check_all) and otherwise recomputes with ``deferred=False``.
    """

    # the grid spans [-1, 1]^3; points outside fall into the boundary cells (the
    # kernels clamp cell indices), which keeps every list a superset of what
    # the reach argument needs, so no rescaling is required
    q = queries.to(torch.float64).contiguous()
    v = verts.to(torch.float64).contiguous()
    f = faces.to(torch.int32).contiguous()
    nq = int(q.shape[0])
    dev = q.device
    lib = _lib.lib()
    face = torch.empty(nq, dtype=torch.int32, device=dev)
    idx = torch.empty((nq, 3), dtype=torch.int32, device=dev)
    w = torch.empty((nq, 3), dtype=torch.float64, device=dev)
    nf = int(f.shape[0])
    ws = _lib.workspace(lib.csg_closest_triangles_workspace_bytes(nf), dev)
    stream = _lib.stream_handle(dev)

    def locate(index, reach, far_dev, far_host):
        if block:
            _lib.check(lib.csg_closest_triangles_block(
                _lib.ptr(q), nq, _lib.ptr(v), _lib.ptr(f), nf, _lib.ptr(index.cell_off),
                _lib.ptr(index.cell_faces), grid_size, int(round(reach * grid_size / 2.0)),
                _lib.ptr(face), _lib.ptr(idx), _lib.ptr(w), _lib.ptr(far_dev), far_host,
                _lib.ptr(ws), ws.numel(), stream), "csg_closest_triangles_block")
        elif _PRUNE:
            _lib.check(lib.csg_closest_triangles_pruned(
                _lib.ptr(q), nq, _lib.ptr(v), _lib.ptr(f), nf, _lib.ptr(index.cell_off),
                _lib.ptr(index.cell_faces), grid_size, reach, _lib.ptr(face), _lib.ptr(idx),
                _lib.ptr(w), _lib.ptr(far_dev), far_host, _lib.ptr(ws), ws.numel(), stream),
                "csg_closest_triangles_pruned")
        else:
            if far_host is None:
                far_dev.zero_()
            _lib.check(lib.csg_closest_triangles(
                _lib.ptr(q), nq, _lib.ptr(v), _lib.ptr(f), _lib.ptr(index.cell_off),
                _lib.ptr(index.cell_faces), grid_size, reach, _lib.ptr(face), _lib.ptr(idx),
                _lib.ptr(w), far_host, _lib.ptr(far_dev), 4, stream), "csg_closest_triangles")

    index_margin = (lambda m: 0) if block else (lambda m: m)
    key = (nf, int(grid_size), index_margin(int(margin)))
    if deferred and key in _CAPACITY:
        capacity = _CAPACITY[key]
        status = torch.zeros(2, dtype=torch.int32, device=dev)
        index = build_grid_index_async(v, f, grid_size, index_margin(margin), capacity,
                                       status[0:1])
        locate(index, margin * 2.0 / grid_size, status[1:], None)
        return face, idx, w, LocateCheck(status, capacity, key)
    far_dev = torch.zeros(1, dtype=torch.int32, device=dev)
    while True:
        index = build_grid_index(v, f, grid_size, index_margin(margin))
        far = ctypes.c_int()
        reach = margin * 2.0 / grid_size
        locate(index, reach, far_dev, ctypes.byref(far))
        if far.value == 0:
            return (face, idx, w, None) if deferred else (face, idx, w)
        if margin >= max_margin:
            raise RuntimeError(f"{far.value} points lie farther than {reach} from the mesh")
        margin += 1


def check_all(checks) -> list[bool]:
    """Confirm LocateChecks (None counts as passed) with one device-to-host copy."""

    live = [c for c in checks if c is not None]
    if not live:
        return [True] * len(checks)
    host = torch.stack([c.status for c in live]).cpu().tolist()
    verdict = iter([c.passed(h) for c, h in zip(live, host)])
    return [True if c is None else next(verdict) for c in checks]
