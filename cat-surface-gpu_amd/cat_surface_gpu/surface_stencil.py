"""Spherical resampling stencils and their device-resident operators (MI355X).

Drop-in for src/cat_surface_gpu/surface_stencil.py of the reference: the
same stencil v1/v2 file format and the same ``SurfaceStencil`` /
``SurfaceStencilDevice`` API.  Every device operator runs in libcsg_hip.so;
topology tables (ascending neighbour CSR, incident-corner CSR, greedy colour
groups, depth-system CSR layout) are built once per face array on the host
and cached on the device, as the reference caches its neighbour tables
(surface_stencil.py:19-33).
"""

from __future__ import annotations

from dataclasses import dataclass, field
import math
from pathlib import Path
import struct
from typing import Sequence

import numpy as np
import torch

from . import _lib, schedule

_HEADER = struct.Struct("<8i")
_MAGIC = 0x46534354
_TOPOLOGY_CACHE: dict[tuple, "Topology"] = {}


# --------------------------------------------------------------------- topology


def build_neighbour_csr(faces: np.ndarray, n_points: int) -> tuple[np.ndarray, np.ndarray]:
    """Unique neighbours of each vertex in ascending order (surface_stencil.py:36-64)."""

    f = np.asarray(faces, dtype=np.int64)
    directed = np.concatenate([f[:, [0, 1]], f[:, [1, 0]], f[:, [1, 2]], f[:, [2, 1]],
                               f[:, [2, 0]], f[:, [0, 2]]])
    directed = np.unique(directed, axis=0)
    counts = np.bincount(directed[:, 0], minlength=n_points)
    off = np.zeros(n_points + 1, np.int32)
    off[1:] = np.cumsum(counts)
    return off, np.ascontiguousarray(directed[:, 1].astype(np.int32))


def build_incident_csr(faces: np.ndarray, n_points: int) -> tuple[np.ndarray, np.ndarray]:
    """face*3+corner per vertex in (corner, face) order: the order of index_add_."""

    f = np.asarray(faces, dtype=np.int64)
    nf = f.shape[0]
    corner_major = f.T.reshape(-1)  # entry e = corner * nf + face
    order = np.argsort(corner_major, kind="stable")
    corner = order // nf
    face = order - corner * nf
    counts = np.bincount(corner_major, minlength=n_points)
    off = np.zeros(n_points + 1, np.int32)
    off[1:] = np.cumsum(counts)
    return off, np.ascontiguousarray((face * 3 + corner).astype(np.int32))


def greedy_colours(off: np.ndarray, nbr: np.ndarray) -> np.ndarray:
    """Smallest colour not used by an already coloured neighbour, in vertex order.

    surface_stencil.py:67-86
    """

    n = off.shape[0] - 1
    colours = np.full(n, -1, dtype=np.int32)
    offl = off.tolist()
    nbrl = nbr.tolist()
    col = colours.tolist()
    for v in range(n):
        used = {col[u] for u in nbrl[offl[v]:offl[v + 1]] if col[u] >= 0}
        c = 0
        while c in used:
            c += 1
        col[v] = c
    return np.asarray(col, dtype=np.int32)


@dataclass
class Topology:
    """Device tables derived from one face array."""

    n_points: int
    faces: torch.Tensor  # int32 [F, 3]
    nbr_off: torch.Tensor
    nbr: torch.Tensor
    inc_off: torch.Tensor
    inc: torch.Tensor
    colour_off: np.ndarray  # host int32 [C + 1]
    colour_verts: torch.Tensor
    sys_row_off: torch.Tensor     # merged (row, col) CSR of the depth system
    sys_cols: torch.Tensor
    sys_contrib_off: torch.Tensor  # per merged entry: its COO contributions
    sys_contrib: torch.Tensor      # COO entry indices, COO order within an entry
    colour_off_dev: torch.Tensor = None
    faces_np: np.ndarray = field(repr=False, default=None)
    # the same system as ELL in colour order: slot k of colour-ordered row r is
    # at k * n + r; ell_src names its merged CSR entry (-1 = padding)
    ell_width: int = 0
    ell_cols: torch.Tensor = None
    ell_src: torch.Tensor = None
    nbr_off_np: np.ndarray = field(repr=False, default=None)
    nbr_np: np.ndarray = field(repr=False, default=None)
    colour_order_np: np.ndarray = field(repr=False, default=None)  # solve order -> vertex
    ell_cols_np: np.ndarray = field(repr=False, default=None)      # [ell_width, n]
    patch_plans: dict = field(repr=False, default_factory=dict)

    @property
    def nnz(self) -> int:
        return int(self.nbr.numel())


@dataclass
class SmoothingPatches:
    """A partition of a mesh for csg_smooth_geometry_patched and
    csg_smooth_values_patched (device arrays)."""

    count: int
    rings: int
    vstart: torch.Tensor
    ring_end: torch.Tensor
    lverts: torch.Tensor
    aoff: torch.Tensor
    astart: torch.Tensor
    ladj: torch.Tensor  # int16 storage of the uint16 local indices
    perm: torch.Tensor  # working order: working index i holds vertex perm[i]
    lds_bytes: int
    largest: int  # local vertices of the largest patch
    value_lds_bytes: int = 0  # LDS of csg_smooth_values_patched: 26 L + 10 E, largest patch


def _morton_order(points: np.ndarray) -> np.ndarray:
    """Vertices in Morton (Z-curve) order of their positions, 10 bits per axis."""

    p = np.asarray(points, dtype=np.float64)
    lo, hi = p.min(0), p.max(0)
    q = np.clip(((p - lo) / np.maximum(hi - lo, 1e-30) * 1023.0).astype(np.int64), 0, 1023)
    code = np.zeros(p.shape[0], np.int64)
    for bit in range(10):
        for d in range(3):
            code |= ((q[:, d] >> bit) & 1) << (3 * bit + d)
    return np.argsort(code, kind="stable")


def _host_points(positions, n: int) -> np.ndarray:
    pos = positions.detach().cpu().numpy() if isinstance(positions, torch.Tensor) else positions
    return np.asarray(pos).reshape(n, 3)


def _patch_layout(topology: Topology, order: np.ndarray, core: int, rings: int):
    """Patches of ``core`` consecutive vertices of ``order`` and the ``rings`` rings
    of their graph neighbourhood: (count, patch, vertex, ring, sizes, vstart) with
    the (patch, vertex, ring) triples sorted by patch, then ring (cores in
    ``order``, each ring by ascending vertex)."""

    import scipy.sparse as sp

    n = topology.n_points
    off = topology.nbr_off_np.astype(np.int64)
    nbr = topology.nbr_np.astype(np.int64)
    count = (n + core - 1) // core
    if count > 65535:
        raise ValueError("too many patches; use a larger core")
    pid = np.empty(n, np.int64)
    pid[order] = np.arange(n) // core
    adj = sp.csr_matrix((np.ones(nbr.size, np.int32), nbr, off), shape=(n, n))
    # ring 0: each patch's core in `order`; ring r + 1: unvisited neighbours of ring r
    ring_pairs = [(pid[order], order)]
    member = sp.csr_matrix((np.ones(n, np.int32), (pid, np.arange(n))), shape=(count, n))
    visited, frontier = member.copy(), member
    for _ in range(rings):
        nxt = frontier @ adj
        nxt.data[:] = 1
        nxt = nxt - nxt.multiply(visited)
        nxt.eliminate_zeros()
        nxt.data[:] = 1
        nxt.sort_indices()
        visited = visited + nxt
        frontier = nxt
        coo = nxt.tocoo()
        ring_pairs.append((coo.row.astype(np.int64), coo.col.astype(np.int64)))
    patch = np.concatenate([rp[0] for rp in ring_pairs])
    vert = np.concatenate([rp[1] for rp in ring_pairs])
    ring = np.concatenate([np.full(rp[0].size, r, np.int64) for r, rp in enumerate(ring_pairs)])
    layout = np.lexsort((np.arange(patch.size), ring, patch))  # stable within a ring
    patch, vert, ring = patch[layout], vert[layout], ring[layout]
    sizes = np.bincount(patch, minlength=count)
    vstart = np.zeros(count + 1, np.int64)
    vstart[1:] = np.cumsum(sizes)
    return count, patch, vert, ring, sizes, vstart


def _local_lookup(patch: np.ndarray, vert: np.ndarray, n: int, vstart: np.ndarray):
    """A function mapping (patch, vertex) arrays to local indices in the patch."""

    key = patch * n + vert
    sort = np.argsort(key, kind="stable")
    sorted_key = key[sort]
    local = np.arange(patch.size) - vstart[patch]

    def find(p, v):
        want = p * n + v
        at = np.searchsorted(sorted_key, want)
        if at.size and (at.max() >= sorted_key.size or np.any(sorted_key[at] != want)):
            raise AssertionError("a patch row reaches outside its patch")
        return local[sort[at]]

    return find


def smoothing_patches(topology: Topology, positions, *, core: int = 256,
                      rings: int = 8) -> SmoothingPatches:
    """Patches of ``core`` vertices (consecutive in Morton order of ``positions``),
    each with the ``rings`` rings of its graph neighbourhood, for iterating the
    geometry smoothing ``rings`` times per launch; cached on the topology per
    (core, rings) (the partition only affects speed, never values)."""

    key = (int(core), int(rings))
    plan = topology.patch_plans.get(key)
    if plan is not None:
        return plan
    if core < 1 or not 1 <= rings <= 64:
        raise ValueError("core must be positive and rings in [1, 64]")
    n = topology.n_points
    off = topology.nbr_off_np.astype(np.int64)
    nbr = topology.nbr_np.astype(np.int64)
    order = _morton_order(_host_points(positions, n))
    count, patch, vert, ring, sizes, vstart = _patch_layout(topology, order, core, rings)
    if sizes.max() >= 65536:
        raise ValueError("a smoothing patch has more than 65535 local vertices")
    per_ring = np.zeros((count, rings + 1), np.int64)
    np.add.at(per_ring, (patch, ring), 1)
    ring_end = np.cumsum(per_ring, axis=1)
    # local adjacency of every vertex in rings < rings, in the mesh's CSR order
    deg = np.where(ring < rings, off[vert + 1] - off[vert], 0)
    src = np.repeat(np.arange(patch.size), deg)
    first = np.repeat(off[vert], deg)
    within = np.arange(src.size) - np.repeat(np.cumsum(deg) - deg, deg)
    nb = nbr[first + within]
    ladj = _local_lookup(patch, vert, n, vstart)(patch[src], nb)
    # per patch CSR offsets relative to the patch's first adjacency entry
    edges = np.bincount(patch, weights=deg, minlength=count).astype(np.int64)
    astart = np.zeros(count + 1, np.int64)
    astart[1:] = np.cumsum(edges)
    aoff = np.zeros(patch.size + count, np.int64)
    cum = np.cumsum(deg) - np.repeat(astart[:-1], sizes)  # end of each vertex's list
    slot = np.arange(patch.size) + patch + 1
    aoff[slot] = cum
    if edges.max() >= 65536:
        raise ValueError("a smoothing patch has more than 65535 adjacency entries")
    lds = 24 * sizes + 2 * (sizes + 1) + 2 * edges
    lds_bytes = int((lds.max() + 15) // 16 * 16)
    vlds = 26 * sizes + 10 * edges + 2
    value_lds_bytes = int((vlds.max() + 15) // 16 * 16)

    def dev(a, dtype=torch.int32):
        return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=topology.nbr.device)

    # working order = Morton order: cores are contiguous, halos lie in nearby cores
    rank = np.empty(n, np.int64)
    rank[order] = np.arange(n)
    plan = SmoothingPatches(
        count=count, rings=int(rings), vstart=dev(vstart), ring_end=dev(ring_end.reshape(-1)),
        lverts=dev(rank[vert]), aoff=dev(aoff), astart=dev(astart), perm=dev(order),
        ladj=dev(ladj.astype(np.uint16).view(np.int16), torch.int16), lds_bytes=lds_bytes,
        largest=int(sizes.max()), value_lds_bytes=value_lds_bytes)
    topology.patch_plans[key] = plan
    return plan


def topology_for(faces: np.ndarray, n_points: int, device: torch.device) -> Topology:
    f = np.ascontiguousarray(np.asarray(faces, dtype=np.int32))
    cache_key = (device.type, device.index, n_points, f.shape, f.tobytes())
    cached = _TOPOLOGY_CACHE.get(cache_key)
    if cached is not None:
        return cached
    nbr_off, nbr = build_neighbour_csr(f, n_points)
    inc_off, inc = build_incident_csr(f, n_points)
    colours = greedy_colours(nbr_off, nbr)
    order = np.argsort(colours, kind="stable")
    ccount = np.bincount(colours)
    colour_off = np.zeros(ccount.size + 1, np.int32)
    colour_off[1:] = np.cumsum(ccount)
    # depth-system layout: COO blocks of rotation_feature.py:144-161, then the alpha
    # diagonal; duplicate (row, col) entries are merged into one CSR entry whose
    # value is the sum of its COO contributions taken in COO order
    i0, i1, i2 = (f[:, k].astype(np.int64) for k in range(3))
    rows = np.concatenate([i0, i1, i2, i0, i1, i1, i2, i0, i2, np.arange(n_points)])
    cols = np.concatenate([i0, i1, i2, i1, i0, i2, i1, i2, i0, np.arange(n_points)])
    flat = rows * n_points + cols
    perm = np.argsort(flat, kind="stable")
    uniq, counts = np.unique(flat[perm], return_counts=True)
    m_rows = uniq // n_points
    m_cols = uniq % n_points
    row_off = np.zeros(n_points + 1, np.int32)
    row_off[1:] = np.cumsum(np.bincount(m_rows, minlength=n_points))
    contrib_off = np.zeros(uniq.size + 1, np.int32)
    contrib_off[1:] = np.cumsum(counts)

    # colour-ordered ELL: row r of the solve order is vertex order[r]; its slots
    # hold the merged CSR entries of that row in CSR (column) order
    lengths = np.diff(row_off)
    width = int(lengths.max()) if lengths.size else 0
    r_len = lengths[order]
    r_start = row_off[:-1][order]
    slot = np.arange(width)[:, None]
    valid = slot < r_len[None, :]
    ell_src = np.where(valid, r_start[None, :] + slot, -1).astype(np.int32)
    ell_cols = np.where(valid, m_cols[np.clip(ell_src, 0, None)], order[None, :]).astype(np.int32)

    def dev(a, dtype=torch.int32):
        return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=device)

    topo = Topology(
        n_points=n_points,
        faces=dev(f),
        nbr_off=dev(nbr_off),
        nbr=dev(nbr),
        inc_off=dev(inc_off),
        inc=dev(inc),
        colour_off=colour_off,
        colour_verts=dev(order.astype(np.int32)),
        sys_row_off=dev(row_off),
        sys_cols=dev(m_cols.astype(np.int32)),
        sys_contrib_off=dev(contrib_off),
        sys_contrib=dev(perm.astype(np.int32)),
        colour_off_dev=dev(colour_off),
        faces_np=f,
        nbr_off_np=np.asarray(nbr_off),
        nbr_np=np.asarray(nbr),
        colour_order_np=order.astype(np.int64),
        ell_cols_np=ell_cols,
        ell_width=width,
        ell_cols=dev(ell_cols.reshape(-1)),
        ell_src=dev(ell_src.reshape(-1)),
    )
    _TOPOLOGY_CACHE[cache_key] = topo
    return topo


class SorPatchesABI(_lib.ctypes.Structure):
    """csg_sor_patches of include/csg.h."""

    _fields_ = [("count", _lib.ctypes.c_int), ("rings", _lib.ctypes.c_int),
                ("max_local", _lib.ctypes.c_int), ("max_rows", _lib.ctypes.c_int)] + [
        (name, _lib.ctypes.c_void_p) for name in
        ("vstart", "ncore", "lverts", "estart", "e_row", "e_local", "cstart", "active",
         "e_lcol")]


@dataclass
class SorPatches:
    """A partition of a mesh for csg_depth_potential_patched (device arrays)."""

    count: int
    rings: int
    max_local: int
    max_rows: int
    arrays: dict  # name -> device tensor

    def abi(self) -> SorPatchesABI:
        a = self.arrays
        return SorPatchesABI(self.count, self.rings, self.max_local, self.max_rows,
                             *[_lib.ptr(a[k]) for k in ("vstart", "ncore", "lverts", "estart",
                                                        "e_row", "e_local", "cstart", "active",
                                                        "e_lcol")])


def sor_patches(topology: Topology, positions, *, core: int) -> SorPatches:
    """Patches of ``core`` vertices (consecutive in Morton order of ``positions``)
    with as many rings as the topology has colours, and their SOR row tables
    (csg_sor_patches); cached on the topology per core size (the partition only
    affects speed, never values)."""

    key = ("sor", int(core))
    plan = topology.patch_plans.get(key)
    if plan is not None:
        return plan
    n = topology.n_points
    colour_off = np.asarray(topology.colour_off, np.int64)
    ncolours = int(colour_off.size - 1)
    rings = ncolours
    order = _morton_order(_host_points(positions, n))
    count, patch, vert, ring, sizes, vstart = _patch_layout(topology, order, int(core), rings)
    find = _local_lookup(patch, vert, n, vstart)
    row_of = np.empty(n, np.int64)
    row_of[topology.colour_order_np] = np.arange(n)
    colour_of = np.searchsorted(colour_off, row_of, side="right") - 1
    local = np.arange(patch.size) - vstart[patch]
    sel = ring < rings
    pe, ve, re_, le = patch[sel], vert[sel], ring[sel], local[sel]
    ce, rowe = colour_of[ve], row_of[ve]
    srt = np.lexsort((le, re_, ce, pe))
    pe, ve, re_, le, ce, rowe = pe[srt], ve[srt], re_[srt], le[srt], ce[srt], rowe[srt]
    rows = np.bincount(pe, minlength=count)
    estart = np.zeros(count + 1, np.int64)
    estart[1:] = np.cumsum(rows)
    per_colour = np.zeros((count, ncolours), np.int64)
    np.add.at(per_colour, (pe, ce), 1)
    cstart = np.zeros((count, ncolours + 1), np.int64)
    cstart[:, 1:] = np.cumsum(per_colour, axis=1)
    active = np.zeros((count, rings), np.int64)
    live = re_ <= rings - 1 - ce  # colour t's entries updated in phase t
    np.add.at(active, (pe[live], ce[live]), 1)
    width = int(topology.ell_width)
    lcol = np.stack([find(pe, topology.ell_cols_np[k][rowe]) for k in range(width)], axis=1)
    if sizes.max() >= 65536:
        raise ValueError("an SOR patch has more than 65535 local vertices")
    ncore = np.bincount(patch[ring == 0], minlength=count)

    def dev(a, dtype=torch.int32):
        return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=topology.nbr.device)

    plan = SorPatches(
        count=count, rings=rings, max_local=int(sizes.max()), max_rows=int(rows.max()),
        arrays={"vstart": dev(vstart), "ncore": dev(ncore), "lverts": dev(vert),
                "estart": dev(estart), "e_row": dev(rowe), "e_local": dev(le),
                "cstart": dev(cstart.reshape(-1)), "active": dev(active.reshape(-1)),
                "e_lcol": dev(lcol.astype(np.uint16).view(np.int16), torch.int16)})
    topology.patch_plans[key] = plan
    return plan


# --------------------------------------------------------------------- file format


@dataclass(frozen=True)
class SurfaceStencil:
    """Store CPU indices, barycentric weights, topology, and unit-sphere points."""

    sphere_points: np.ndarray
    faces: np.ndarray
    surface_indices: np.ndarray
    surface_weights: np.ndarray
    sheet_indices: np.ndarray
    sheet_weights: np.ndarray
    nx: int
    ny: int
    source_points: int
    unit_sphere_points: np.ndarray | None = None

    @classmethod
    def from_file(cls, path: str | Path) -> "SurfaceStencil":
        """Read a versioned stencil file (surface_stencil.py:127-191)."""

        with Path(path).open("rb") as stream:
            header = stream.read(_HEADER.size)
            if len(header) != _HEADER.size:
                raise ValueError("Stencil file is too short")
            magic, version, n_points, n_tri, n_sheet, nx, ny, source_points = _HEADER.unpack(header)
            if magic != _MAGIC or version not in (1, 2):
                raise ValueError("Stencil file magic or version is invalid")
            sphere_points = np.fromfile(stream, dtype="<f8", count=3 * n_points)
            faces = np.fromfile(stream, dtype="<i4", count=3 * n_tri)
            surface_indices = np.fromfile(stream, dtype="<i4", count=3 * n_points)
            surface_weights = np.fromfile(stream, dtype="<f8", count=3 * n_points)
            sheet_indices = np.fromfile(stream, dtype="<i4", count=3 * n_sheet)
            sheet_weights = np.fromfile(stream, dtype="<f8", count=3 * n_sheet)
            unit = np.fromfile(stream, dtype="<f4", count=3 * source_points) if version == 2 else None
        arrays = (sphere_points, faces, surface_indices, surface_weights, sheet_indices,
                  sheet_weights)
        expected = (3 * n_points, 3 * n_tri, 3 * n_points, 3 * n_points, 3 * n_sheet, 3 * n_sheet)
        if any(a.size != e for a, e in zip(arrays, expected)):
            raise ValueError("Stencil file length is invalid")
        if unit is not None and unit.size != 3 * source_points:
            raise ValueError("Stencil unit sphere length is invalid")
        return cls(
            sphere_points=sphere_points.reshape(n_points, 3),
            faces=faces.reshape(n_tri, 3),
            surface_indices=surface_indices.reshape(n_points, 3),
            surface_weights=surface_weights.reshape(n_points, 3),
            sheet_indices=sheet_indices.reshape(ny, nx, 3),
            sheet_weights=sheet_weights.reshape(ny, nx, 3),
            nx=nx,
            ny=ny,
            source_points=source_points,
            unit_sphere_points=None if unit is None else unit.reshape(source_points, 3),
        )

    def to_file(self, path: str | Path) -> None:
        """Write the v2 layout of native/cat_surface_stencil_builder.c:428-461."""

        version = 2 if self.unit_sphere_points is not None else 1
        with Path(path).open("wb") as stream:
            stream.write(_HEADER.pack(_MAGIC, version, self.sphere_points.shape[0],
                                      self.faces.shape[0], self.nx * self.ny, self.nx, self.ny,
                                      self.source_points))
            np.asarray(self.sphere_points, "<f8").tofile(stream)
            np.asarray(self.faces, "<i4").tofile(stream)
            np.asarray(self.surface_indices, "<i4").tofile(stream)
            np.asarray(self.surface_weights, "<f8").tofile(stream)
            np.asarray(self.sheet_indices, "<i4").tofile(stream)
            np.asarray(self.sheet_weights, "<f8").tofile(stream)
            if self.unit_sphere_points is not None:
                np.asarray(self.unit_sphere_points, "<f4").tofile(stream)

    def to(self, device: str | torch.device, *, geometry_dtype: torch.dtype = torch.float32,
           weight_dtype: torch.dtype = torch.float64) -> "SurfaceStencilDevice":
        """Upload the stencil and its cached topology tables (surface_stencil.py:193-261)."""

        from .dartel_grid import resolve_device

        target = resolve_device(device)
        topo = topology_for(self.faces, int(self.sphere_points.shape[0]), target)

        def dev(a, dtype):
            return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=target)

        return SurfaceStencilDevice(
            sphere_points=dev(self.sphere_points, geometry_dtype),
            faces=topo.faces,
            surface_indices=dev(self.surface_indices, torch.int32),
            surface_weights=dev(self.surface_weights, torch.float64),
            sheet_indices=dev(self.sheet_indices, torch.int32),
            sheet_weights=dev(self.sheet_weights, torch.float64),
            nx=self.nx,
            ny=self.ny,
            source_points=self.source_points,
            unit_sphere_points=(None if self.unit_sphere_points is None
                                else dev(self.unit_sphere_points, torch.float32)),
            topology=topo,
            sphere_points64=dev(self.sphere_points, torch.float64),
        )


# --------------------------------------------------------------------- device ops


def heat_parameters_from_edge_mean(edge_mean: float, fwhm: float) -> tuple[int, float]:
    """iterations = ⌈fwhm²·0.541011/ē²⌉, σ = fwhm·0.7355345/√iterations.

    surface_stencil.py:359-370
    """

    raw = float(fwhm) * float(fwhm) * 0.541011 / (edge_mean * edge_mean)
    iterations = max(1, int(np.ceil(raw)))
    sigma = float(fwhm) * 0.7355345 / float(iterations ** 0.5)
    return iterations, sigma


@dataclass
class SurfaceStencilDevice:
    """Device-resident stencil and topology tables for repeated use."""

    sphere_points: torch.Tensor
    faces: torch.Tensor
    surface_indices: torch.Tensor
    surface_weights: torch.Tensor
    sheet_indices: torch.Tensor
    sheet_weights: torch.Tensor
    nx: int
    ny: int
    source_points: int
    unit_sphere_points: torch.Tensor | None = None
    topology: Topology | None = None
    sphere_points64: torch.Tensor | None = None  # float64 copy for index building

    @property
    def device(self) -> torch.device:
        return self.sphere_points.device

    def _stream(self) -> int:
        return _lib.stream_handle(self.device)

    def _call(self, name: str, *args) -> None:
        _lib.check(getattr(_lib.lib(), name)(*args), name)

    # ---- resampling and sheets

    def resample_vertices(self, vertices: torch.Tensor) -> torch.Tensor:
        """Resample source vertices onto the stencil sphere (surface_stencil.py:282-306)."""

        value = torch.as_tensor(vertices, device=self.device)
        if value.ndim != 2 or tuple(value.shape) != (self.source_points, 3):
            raise ValueError(f"vertices shape must be {(self.source_points, 3)}, "
                             f"got {tuple(value.shape)}")
        if value.dtype not in (torch.float32, torch.float64):
            value = value.to(torch.float32)
        value = value.contiguous()
        n = int(self.surface_indices.shape[0])
        out = torch.empty((n, 3), dtype=value.dtype, device=self.device)
        self._call("csg_resample_vertices", _lib.ptr(value), _lib.dtype_code(value.dtype),
                   _lib.ptr(self.surface_indices), _lib.ptr(self.surface_weights), n,
                   _lib.ptr(out), self._stream())
        return out

    def map_values_to_sheet(self, values: torch.Tensor) -> torch.Tensor:
        """sheet = Σ_c values[idx]·w (surface_stencil.py:308-317)."""

        value = torch.as_tensor(values, dtype=torch.float64, device=self.device).reshape(-1)
        if value.numel() != self.sphere_points.shape[0]:
            raise ValueError(f"values must contain {self.sphere_points.shape[0]} elements, "
                             f"got {value.numel()}")
        value = value.contiguous()
        out = torch.empty((self.ny, self.nx), dtype=torch.float64, device=self.device)
        self._call("csg_sheet_gather", _lib.ptr(value), value.numel(), 1,
                   _lib.ptr(self.sheet_indices), _lib.ptr(self.sheet_weights),
                   self.nx * self.ny, _lib.ptr(out), self._stream())
        return out

    # ---- heat kernel

    def edge_mean_distance(self, vertices: torch.Tensor) -> float:
        """Mean face-edge length (surface_stencil.py:349-357); synchronises."""

        return edge_means([self], [vertices])[0]

    def _heat_parameters(self, vertices: torch.Tensor, fwhm: float) -> tuple[int, float]:
        return heat_parameters_from_edge_mean(self.edge_mean_distance(vertices), fwhm)

    def smooth_geometry(self, vertices: torch.Tensor, fwhm: float,
                        iterations: int | None = None) -> torch.Tensor:
        """Iterative heat-kernel smoothing of geometry (surface_stencil.py:372-388)."""

        v = torch.as_tensor(vertices, dtype=torch.float32, device=self.device).contiguous()
        if iterations is None:
            iterations, _ = self._heat_parameters(v, fwhm)
        topo = self.topology
        n = int(v.shape[0])
        out = torch.empty_like(v)
        ws = torch.empty_like(v)
        self._call("csg_smooth_geometry", _lib.ptr(v), _lib.ptr(topo.nbr_off), _lib.ptr(topo.nbr),
                   n, int(iterations), _lib.ptr(out), _lib.ptr(ws), ws.numel() * 4,
                   self._stream())
        return out

    def _smooth_values_batch(self, values: torch.Tensor, geometry: torch.Tensor,
                             params: list[tuple[int, float]], squared_norm: bool,
                             mesh_of: Sequence[int] | None = None) -> torch.Tensor:
        """Rows of ``values`` smoothed with (iterations, sigma) ``params``.

        ``geometry`` is one mesh [n, 3] or a stack [meshes, n, 3] of meshes of
        this stencil's topology; row b uses mesh ``mesh_of[b]`` (default 0).
        """

        topo = self.topology
        geom = geometry.to(torch.float32).contiguous()
        meshes = 1 if geom.ndim == 2 else int(geom.shape[0])
        n = int(geom.shape[-2])
        nb = len(params)
        vals = values.to(torch.float64).contiguous()
        out = torch.empty((nb, n), dtype=torch.float64, device=self.device)
        ws = _lib.workspace(_lib.lib().csg_smooth_values_workspace_bytes(n, topo.nnz, nb),
                            self.device)
        sig = _lib.doubles([p[1] for p in params])
        its = (_lib.ctypes.c_int * nb)(*[int(p[0]) for p in params])
        mesh = None if mesh_of is None else (_lib.ctypes.c_int * nb)(*[int(m) for m in mesh_of])
        if _VALUE_PATCHES is not None and topo.nbr_off_np is not None:
            pp = smoothing_patches(topo, self.sphere_points, core=_VALUE_PATCHES[0],
                                   rings=_VALUE_PATCHES[1])
            if pp.value_lds_bytes <= 160 * 1024:
                self._call("csg_smooth_values_patched", _lib.ptr(vals), n, nb, _lib.ptr(geom),
                           meshes, mesh, _lib.ptr(topo.nbr_off), _lib.ptr(topo.nbr), topo.nnz,
                           sig, its, int(squared_norm), pp.count, pp.rings, _lib.ptr(pp.perm),
                           _lib.ptr(pp.vstart), _lib.ptr(pp.ring_end), _lib.ptr(pp.lverts),
                           _lib.ptr(pp.aoff), _lib.ptr(pp.astart), _lib.ptr(pp.ladj),
                           pp.value_lds_bytes, _lib.ptr(out), _lib.ptr(ws), ws.numel(),
                           self._stream())
                return out
        self._call("csg_smooth_values", _lib.ptr(vals), n, nb, _lib.ptr(geom), meshes, mesh,
                   _lib.ptr(topo.nbr_off), _lib.ptr(topo.nbr), topo.nnz, sig, its,
                   int(squared_norm), _lib.ptr(out), _lib.ptr(ws), ws.numel(), self._stream())
        return out

    def _smooth_values_batch_dev(self, values: torch.Tensor, geometry: torch.Tensor,
                                 iters: torch.Tensor, sigma: torch.Tensor, squared_norm: bool,
                                 mesh_of: Sequence[int] | None = None) -> torch.Tensor:
        """_smooth_values_batch with the rows' iterations and widths in device
        arrays (iters int32 [nb], sigma float64 [nb]); launches to
        schedule.launch_bound."""

        topo = self.topology
        geom = geometry.to(torch.float32).contiguous()
        meshes = 1 if geom.ndim == 2 else int(geom.shape[0])
        n = int(geom.shape[-2])
        nb = int(iters.numel())
        its = iters.to(torch.int32).contiguous()
        sig = sigma.to(torch.float64).contiguous()
        bound = schedule.launch_bound([its], "value smoothing iterations")
        vals = values.to(torch.float64).contiguous()
        out = torch.empty((nb, n), dtype=torch.float64, device=self.device)
        ws = _lib.workspace(_lib.lib().csg_smooth_values_workspace_bytes(n, topo.nnz, nb),
                            self.device)
        mesh = None if mesh_of is None else (_lib.ctypes.c_int * nb)(*[int(m) for m in mesh_of])
        if _VALUE_PATCHES is not None and topo.nbr_off_np is not None:
            pp = smoothing_patches(topo, self.sphere_points, core=_VALUE_PATCHES[0],
                                   rings=_VALUE_PATCHES[1])
            if pp.value_lds_bytes <= 160 * 1024:
                self._call("csg_smooth_values_patched_dev", _lib.ptr(vals), n, nb,
                           _lib.ptr(geom), meshes, mesh, _lib.ptr(topo.nbr_off),
                           _lib.ptr(topo.nbr), topo.nnz, _lib.ptr(sig), _lib.ptr(its),
                           int(bound), int(squared_norm), pp.count, pp.rings, _lib.ptr(pp.perm),
                           _lib.ptr(pp.vstart), _lib.ptr(pp.ring_end), _lib.ptr(pp.lverts),
                           _lib.ptr(pp.aoff), _lib.ptr(pp.astart), _lib.ptr(pp.ladj),
                           pp.value_lds_bytes, _lib.ptr(out), _lib.ptr(ws), ws.numel(),
                           self._stream())
                return out
        self._call("csg_smooth_values_dev", _lib.ptr(vals), n, nb, _lib.ptr(geom), meshes, mesh,
                   _lib.ptr(topo.nbr_off), _lib.ptr(topo.nbr), topo.nnz, _lib.ptr(sig),
                   _lib.ptr(its), int(bound), int(squared_norm), _lib.ptr(out), _lib.ptr(ws),
                   ws.numel(), self._stream())
        return out

    def smooth_values(self, values: torch.Tensor, vertices: torch.Tensor,
                      fwhm: float) -> torch.Tensor:
        """Iterative heat-kernel smoothing of values (surface_stencil.py:390-412)."""

        geom = torch.as_tensor(vertices, device=self.device)
        params = [self._heat_parameters(geom, fwhm)]
        vals = torch.as_tensor(values, dtype=torch.float64, device=self.device).reshape(1, -1)
        return self._smooth_values_batch(vals, geom, params, False)[0]

    def smooth_values_many(self, values: torch.Tensor, vertices: torch.Tensor,
                           fwhms: tuple[float, ...]) -> tuple[torch.Tensor, ...]:
        """Several fwhms at once (surface_stencil.py:414-476)."""

        if not fwhms:
            return ()
        base = torch.as_tensor(values, dtype=torch.float64, device=self.device)
        if base.ndim != 1 or base.numel() != self.sphere_points.shape[0]:
            raise ValueError("smooth_values_many values must have shape [sphere_points]")
        geom = torch.as_tensor(vertices, device=self.device)
        if tuple(geom.shape) != tuple(base.shape) + (3,):
            raise ValueError("smooth_values_many vertices must have shape [sphere_points, 3]")
        em = self.edge_mean_distance(geom)
        params = [heat_parameters_from_edge_mean(em, float(f)) for f in fwhms]
        rows = base.reshape(1, -1).expand(len(fwhms), -1)
        out = self._smooth_values_batch(rows, geom, params, True)
        return tuple(out[i] for i in range(len(fwhms)))

    def vertex_normals(self, vertices: torch.Tensor) -> torch.Tensor:
        """Angle-weighted vertex normals (surface_stencil.py:478-505)."""

        v = torch.as_tensor(vertices, dtype=torch.float32, device=self.device).contiguous()
        out = torch.empty_like(v)
        topo = self.topology
        self._call("csg_vertex_normals", _lib.ptr(v), _lib.ptr(topo.faces),
                   _lib.ptr(topo.inc_off), _lib.ptr(topo.inc), int(v.shape[0]), _lib.ptr(out),
                   self._stream())
        return out

    # ---- curvature sheets

    def curvature_type5(self, vertices: torch.Tensor, fwhm: float) -> torch.Tensor:
        """Depth along fwhm-50 smoothed normals, smoothed and normalised (:507-519)."""

        geometry = torch.as_tensor(vertices, dtype=torch.float32, device=self.device).contiguous()
        smoothed = self.smooth_geometry(geometry, 50.0)
        normals = self.vertex_normals(smoothed)
        base = ((smoothed - geometry) * normals).sum(dim=-1).to(torch.float64)
        values = self.smooth_values(base, geometry, fwhm)
        return _minmax(values)

    def curvature_type5_to_sheet_many(self, vertices: torch.Tensor, fwhms: tuple[float, ...],
                                      out_dtype: torch.dtype = torch.float64,
                                      ) -> tuple[torch.Tensor, ...]:
        """One fused call for all fwhms (surface_stencil.py:521-550)."""

        if not fwhms:
            return ()
        geometry = torch.as_tensor(vertices, dtype=torch.float32, device=self.device).contiguous()
        em = self.edge_mean_distance(geometry)
        geom_iters, _ = heat_parameters_from_edge_mean(em, 50.0)
        params = [heat_parameters_from_edge_mean(em, float(f)) for f in fwhms]
        out = self._type5(geometry, geom_iters, params, out_dtype)
        return tuple(out[i] for i in range(len(fwhms)))

    def _type5(self, geometry: torch.Tensor, geom_iters: int, params, out_dtype,
               smoothed: torch.Tensor | None = None) -> torch.Tensor:
        topo = self.topology
        n = int(geometry.shape[0])
        nb = len(params)
        m = self.nx * self.ny
        out = torch.empty((nb, self.ny, self.nx), dtype=out_dtype, device=self.device)
        ws = _lib.workspace(_lib.lib().csg_type5_workspace_bytes(n, topo.nnz, nb, m), self.device)
        sig = _lib.doubles([p[1] for p in params])
        its = (_lib.ctypes.c_int * nb)(*[int(p[0]) for p in params])
        self._call("csg_type5_sheets", _lib.ptr(geometry), n, _lib.ptr(topo.faces),
                   _lib.ptr(topo.nbr_off), _lib.ptr(topo.nbr), topo.nnz, _lib.ptr(topo.inc_off),
                   _lib.ptr(topo.inc), int(geom_iters), _lib.ptr(smoothed), nb, sig, its,
                   _lib.ptr(self.sheet_indices), _lib.ptr(self.sheet_weights), m,
                   _lib.dtype_code(out_dtype), _lib.ptr(out), _lib.ptr(ws), ws.numel(),
                   self._stream())
        return out

    def curvature_type2(self, vertices: torch.Tensor, fwhm: float) -> torch.Tensor:
        """Shape-operator curvature sqrt((k1^2 + k2^2) / 2), smoothed and
        normalised (:552-622)."""

        geometry = torch.as_tensor(vertices, dtype=torch.float32, device=self.device).contiguous()
        normals = self.vertex_normals(geometry)
        topo = self.topology
        raw = torch.empty(int(geometry.shape[0]), dtype=torch.float64, device=self.device)
        self._call("csg_curvature_type2", _lib.ptr(geometry), _lib.ptr(normals),
                   int(geometry.shape[0]), _lib.ptr(topo.nbr_off), _lib.ptr(topo.nbr),
                   _lib.ptr(raw), self._stream())
        return _minmax(self.smooth_values(raw, geometry, fwhm))

    def curvature_to_sheet(self, vertices: torch.Tensor, fwhm: float, *, curvtype: int = 5,
                           out_dtype: torch.dtype = torch.float64) -> torch.Tensor:
        """Supported CAT curvature mapped to a normalised sheet (:624-644): one
        fused call (geometry smoothing for type 5, curvature, value smoothing,
        normalisation, sheet gather, normalisation)."""

        if int(curvtype) not in (2, 5):
            raise NotImplementedError("only upstream curvtype 2 and 5 exist")
        geometry = torch.as_tensor(vertices, dtype=torch.float32, device=self.device).contiguous()
        em = self.edge_mean_distance(geometry)
        geom_iters = heat_parameters_from_edge_mean(em, 50.0)[0] if int(curvtype) == 5 else 0
        iterations, sigma = heat_parameters_from_edge_mean(em, float(fwhm))
        topo = self.topology
        n = int(geometry.shape[0])
        m = self.nx * self.ny
        out = torch.empty((self.ny, self.nx), dtype=out_dtype, device=self.device)
        ws = _lib.workspace(_lib.lib().csg_curvature_sheet_workspace_bytes(n, topo.nnz, m),
                            self.device)
        self._call("csg_curvature_sheet", _lib.ptr(geometry), n, _lib.ptr(topo.faces),
                   _lib.ptr(topo.nbr_off), _lib.ptr(topo.nbr), topo.nnz, _lib.ptr(topo.inc_off),
                   _lib.ptr(topo.inc), int(curvtype), int(geom_iters), float(sigma),
                   int(iterations), _lib.ptr(self.sheet_indices), _lib.ptr(self.sheet_weights), m,
                   _lib.dtype_code(out_dtype), _lib.ptr(out), _lib.ptr(ws), ws.numel(),
                   self._stream())
        return out

    def curvature_sheets(self, vertices: torch.Tensor, fwhms: Sequence[float],
                         curvtypes: Sequence[int],
                         out_dtype: torch.dtype = torch.float64) -> torch.Tensor:
        """[steps, ny, nx] maps of one surface by the reference's rule
        (surface_warp.py:427-448): several steps that are all type 5 share one
        fused call with sum-of-squares heat weights (curvature_type5_to_sheet_many);
        otherwise every step is its own curvature_to_sheet."""

        types = tuple(int(t) for t in curvtypes)
        if len(types) != len(fwhms):
            raise ValueError("one curvtype per fwhm is required")
        if len(types) > 1 and all(t == 5 for t in types):
            return torch.stack(self.curvature_type5_to_sheet_many(
                vertices, tuple(float(f) for f in fwhms), out_dtype=out_dtype))
        return torch.stack([self.curvature_to_sheet(vertices, float(f), curvtype=t,
                                                    out_dtype=out_dtype)
                            for f, t in zip(fwhms, types)])


# ---- heat-kernel parameters on the device

_POW_ROOTS: dict[str, tuple] = {}
_POW_ROOT_LIMIT = 1 << 20


def _pow_root_table(device: torch.device):
    """Counts k <= 2^20 whose k ** 0.5 (the host's libm pow, as
    heat_parameters_from_edge_mean evaluates it) differs from the correctly
    rounded square root, with those pow values, on the device (built once)."""

    key = str(device)
    hit = _POW_ROOTS.get(key)
    if hit is None:
        ks, vs = [], []
        for k in range(1, _POW_ROOT_LIMIT + 1):
            r = k ** 0.5
            if r != math.sqrt(k):
                ks.append(k)
                vs.append(r)
        hit = (torch.tensor(ks, dtype=torch.int32, device=device),
               torch.tensor(vs, dtype=torch.float64, device=device), len(ks))
        _POW_ROOTS[key] = hit
    return hit


@dataclass
class EdgeHeat:
    """One mesh's heat-kernel parameters for a list of fwhms, on the device
    (csg_heat_parameters): iterations [k] int32, sigma [k] float64 and the mean
    face-edge length [1] float64 (surface_stencil.py:349-370)."""

    fwhms: tuple
    iters: torch.Tensor
    sigma: torch.Tensor
    edge_mean: torch.Tensor

    def host(self) -> list[tuple[int, float]]:
        """(iterations, sigma) per fwhm (synchronises)."""

        its = self.iters.cpu().tolist()
        sig = self.sigma.cpu().tolist()
        return list(zip(its, sig))


def edge_heat(stencils, geometries, fwhm_lists) -> list[EdgeHeat]:
    """Heat parameters of several meshes for their fwhm lists, computed and kept
    on the device (no synchronisation)."""

    out = []
    for st, g, fws in zip(stencils, geometries, fwhm_lists):
        v = torch.as_tensor(g, dtype=torch.float32, device=st.device).contiguous()
        nf = int(st.faces.shape[0])
        fws = tuple(float(f) for f in fws)
        ws = _lib.workspace(_lib.lib().csg_edge_mean_workspace_bytes(nf), st.device)
        iters = torch.empty(len(fws), dtype=torch.int32, device=st.device)
        sigma = torch.empty(len(fws), dtype=torch.float64, device=st.device)
        em = torch.empty(1, dtype=torch.float64, device=st.device)
        rn, rv, nr = _pow_root_table(st.device)
        _lib.check(_lib.lib().csg_heat_parameters(
            _lib.ptr(v), _lib.ptr(st.faces), nf, _lib.doubles(fws), len(fws), _lib.ptr(rn),
            _lib.ptr(rv), nr, _lib.ptr(iters), _lib.ptr(sigma), _lib.ptr(em), _lib.ptr(ws),
            ws.numel(), st._stream()), "csg_heat_parameters")
        out.append(EdgeHeat(fws, iters, sigma, em))
    return out


def heat_from_edge_mean(heat: EdgeHeat, fwhms) -> EdgeHeat:
    """Parameters for other fwhms from an EdgeHeat's edge mean (on the device)."""

    dev = heat.edge_mean.device
    fws = tuple(float(f) for f in fwhms)
    iters = torch.empty(len(fws), dtype=torch.int32, device=dev)
    sigma = torch.empty(len(fws), dtype=torch.float64, device=dev)
    rn, rv, nr = _pow_root_table(dev)
    _lib.check(_lib.lib().csg_heat_parameters_from_edge_mean(
        _lib.ptr(heat.edge_mean), _lib.doubles(fws), len(fws), _lib.ptr(rn), _lib.ptr(rv), nr,
        _lib.ptr(iters), _lib.ptr(sigma), _lib.stream_handle(dev)),
        "csg_heat_parameters_from_edge_mean")
    return EdgeHeat(fws, iters, sigma, heat.edge_mean)


def edge_means(stencils, geometries) -> list[float]:
    """edge_mean_distance of several meshes with a single host synchronisation
    (surface_stencil.py:349-357): each mesh's edge-sum partials added in order on
    the device (the edge mean csg_heat_parameters uses)."""

    heats = edge_heat(stencils, geometries, [()] * len(stencils))
    if not heats:
        return []
    return torch.cat([h.edge_mean for h in heats]).cpu().tolist()


# Patch shape (core vertices, rings = iterations per launch) of the batched
# geometry smoothing; None: one launch per iteration.  (512, 4) measured best on
# the 61,442-vertex step (46.6 -> 46.1 ms against (320, 4); 448: 46.1-46.4, 640: 46.4).
_GEOMETRY_PATCHES: tuple[int, int] | None = (512, 4)
# value smoothing: (core, rings) patches iterated in LDS, or None for one launch
# per iteration; identical values
_VALUE_PATCHES: tuple[int, int] | None = (256, 4)


def set_value_smoothing(patches: tuple[int, int] | None) -> tuple[int, int] | None:
    """Select the value smoothing's schedule: (core, rings) patches iterated in
    LDS (csg_smooth_values_patched), or None for one launch per iteration
    (csg_smooth_values).  Both give identical values.  Returns the previous
    setting."""

    global _VALUE_PATCHES
    previous = _VALUE_PATCHES
    if patches is not None:
        core, rings = (int(v) for v in patches)
        if core < 1 or not 1 <= rings <= 64:
            raise ValueError("patches must be (core >= 1, 1 <= rings <= 64)")
        patches = (core, rings)
    _VALUE_PATCHES = patches
    return previous


def set_geometry_smoothing(patches: tuple[int, int] | None) -> tuple[int, int] | None:
    """Select the batched geometry smoothing's schedule: (core, rings) patches
    iterated in LDS (csg_smooth_geometry_patched), or None for one launch per
    iteration (csg_smooth_geometry_batch).  Both give identical values.
    Returns the previous setting."""

    global _GEOMETRY_PATCHES
    previous = _GEOMETRY_PATCHES
    if patches is not None:
        core, rings = (int(v) for v in patches)
        if core < 1 or not 1 <= rings <= 64:
            raise ValueError("patches must be (core >= 1, 1 <= rings <= 64)")
        patches = (core, rings)
    _GEOMETRY_PATCHES = patches
    return previous


def smooth_geometry_many(stencils, geometries, iterations) -> list[torch.Tensor]:
    """SurfaceStencilDevice.smooth_geometry of several meshes with given
    iteration counts; meshes that share a topology are smoothed in the same
    launches (csg_smooth_geometry_patched, or csg_smooth_geometry_batch when
    patches are off).  Results equal one call per mesh.  ``iterations`` holds
    host ints, or one-element device int32 tensors (EdgeHeat.iters slices):
    then the _dev entries read the counts on the device and the launches run to
    schedule.launch_bound."""

    device_counts = len(iterations) > 0 and isinstance(iterations[0], torch.Tensor)
    bound = (schedule.launch_bound(list(iterations), "geometry smoothing iterations")
             if device_counts else 0)
    out: list[torch.Tensor | None] = [None] * len(stencils)
    groups: dict[int, list[int]] = {}
    for i, st in enumerate(stencils):
        groups.setdefault(id(st.topology), []).append(i)
    for members in groups.values():
        for start in range(0, len(members), 32):
            chunk = members[start:start + 32]
            topo = stencils[chunk[0]].topology
            dev = stencils[chunk[0]].device
            v = torch.stack([torch.as_tensor(geometries[i], dtype=torch.float32, device=dev)
                             for i in chunk]).contiguous()
            n = int(v.shape[1])
            res = torch.empty_like(v)
            ws = _lib.workspace(_lib.lib().csg_smooth_geometry_batch_workspace_bytes(
                n, len(chunk)), dev)
            if device_counts:
                its_dev = torch.cat([iterations[i].reshape(1) for i in chunk]).to(
                    torch.int32).contiguous()
                _smooth_geometry_dev(stencils[chunk[0]], v, its_dev, bound, res, ws)
                for k, i in enumerate(chunk):
                    out[i] = res[k]
                continue
            its = (_lib.ctypes.c_int * len(chunk))(*[int(iterations[i]) for i in chunk])
            if _GEOMETRY_PATCHES is not None and topo.nbr_off_np is not None:
                # the partition follows the stencil's sphere (compact patches)
                pp = smoothing_patches(topo, stencils[chunk[0]].sphere_points,
                                       core=_GEOMETRY_PATCHES[0], rings=_GEOMETRY_PATCHES[1])
                _lib.check(_lib.lib().csg_smooth_geometry_patched(
                    _lib.ptr(v), len(chunk), n, its, pp.count, pp.rings, _lib.ptr(pp.perm),
                    _lib.ptr(pp.vstart),
                    _lib.ptr(pp.ring_end), _lib.ptr(pp.lverts), _lib.ptr(pp.aoff),
                    _lib.ptr(pp.astart), _lib.ptr(pp.ladj), pp.lds_bytes, _lib.ptr(res),
                    _lib.ptr(ws), ws.numel(), stencils[chunk[0]]._stream()),
                    "csg_smooth_geometry_patched")
            else:
                _lib.check(_lib.lib().csg_smooth_geometry_batch(
                    _lib.ptr(v), len(chunk), _lib.ptr(topo.nbr_off), _lib.ptr(topo.nbr), n, its,
                    _lib.ptr(res), _lib.ptr(ws), ws.numel(), stencils[chunk[0]]._stream()),
                    "csg_smooth_geometry_batch")
            for k, i in enumerate(chunk):
                out[i] = res[k]
    return out


def _smooth_geometry_dev(st, v: torch.Tensor, its_dev: torch.Tensor, bound: int,
                         res: torch.Tensor, ws: torch.Tensor) -> None:
    topo = st.topology
    batch, n = int(v.shape[0]), int(v.shape[1])
    if _GEOMETRY_PATCHES is not None and topo.nbr_off_np is not None:
        pp = smoothing_patches(topo, st.sphere_points, core=_GEOMETRY_PATCHES[0],
                               rings=_GEOMETRY_PATCHES[1])
        _lib.check(_lib.lib().csg_smooth_geometry_patched_dev(
            _lib.ptr(v), batch, n, _lib.ptr(its_dev), int(bound), pp.count, pp.rings,
            _lib.ptr(pp.perm), _lib.ptr(pp.vstart), _lib.ptr(pp.ring_end), _lib.ptr(pp.lverts),
            _lib.ptr(pp.aoff), _lib.ptr(pp.astart), _lib.ptr(pp.ladj), pp.lds_bytes,
            _lib.ptr(res), _lib.ptr(ws), ws.numel(), st._stream()),
            "csg_smooth_geometry_patched_dev")
        return
    _lib.check(_lib.lib().csg_smooth_geometry_batch_dev(
        _lib.ptr(v), batch, _lib.ptr(topo.nbr_off), _lib.ptr(topo.nbr), n, _lib.ptr(its_dev),
        int(bound), _lib.ptr(res), _lib.ptr(ws), ws.numel(), st._stream()),
        "csg_smooth_geometry_batch_dev")


def prepare_type5(stencils, geometries) -> tuple[list, list, list]:
    """The fwhm-independent part of curvature_type5_to_sheet_many for several
    (stencil, resampled geometry) pairs: float32 geometries, their device heat
    parameters (EdgeHeat, fwhm 50) and the fwhm-50 smoothed geometries (shared
    launches).
    Pairs with different fwhm lists can share it (type5_sheets_many's
    ``prepared``)."""

    geoms = [torch.as_tensor(g, dtype=torch.float32, device=st.device).contiguous()
             for st, g in zip(stencils, geometries)]
    heats = edge_heat(stencils, geoms, [(50.0,)] * len(geoms))
    smoothed = smooth_geometry_many(stencils, geoms, [h.iters[0:1] for h in heats])
    return geoms, heats, smoothed


def type5_sheets_many(stencils, geometries, fwhms, out_dtype=torch.float64, *,
                      prepared=None) -> list[torch.Tensor]:
    """curvature_type5_to_sheet_many for several (stencil, resampled geometry)
    pairs: one host synchronisation for all edge means and the fwhm-50 geometry
    smoothing of every mesh in shared launches (or ``prepared``, prepare_type5's
    result for these pairs).  Returns [len(fwhms), ny, nx] sheets per pair,
    equal to the per-pair calls."""

    fw = tuple(float(f) for f in fwhms)
    geoms, heats, smoothed = (prepared if prepared is not None
                              else prepare_type5(stencils, geometries))
    out: list[torch.Tensor | None] = [None] * len(stencils)
    if not fw:
        return [torch.empty((0, st.ny, st.nx), dtype=out_dtype, device=st.device)
                for st in stencils]
    # sides of one topology and sheet shape share one value-smoothing loop
    groups: dict[tuple, list[int]] = {}
    for i, st in enumerate(stencils):
        groups.setdefault((id(st.topology), st.nx, st.ny, str(st.device)), []).append(i)
    per_call = max(1, 32 // len(fw))
    for members in groups.values():
        for start in range(0, len(members), per_call):
            chunk = members[start:start + per_call]
            out_chunk = _type5_batch([stencils[i] for i in chunk], [geoms[i] for i in chunk],
                                     [smoothed[i] for i in chunk], [heats[i] for i in chunk], fw,
                                     out_dtype)
            for k, i in enumerate(chunk):
                out[i] = out_chunk[k]
    return out


def _type5_batch(stencils, geoms, smoothed, heats, fw, out_dtype) -> torch.Tensor:
    """csg_type5_sheets_batch_dev over sides of one topology and sheet shape, the
    rows' heat parameters computed on the device from each side's edge mean."""

    st0 = stencils[0]
    topo = st0.topology
    sides, nfw = len(stencils), len(fw)
    n = int(geoms[0].shape[0])
    m = st0.nx * st0.ny
    g = torch.stack(geoms).contiguous()
    sm = torch.stack(smoothed).contiguous()
    rows = [heat_from_edge_mean(h, fw) for h in heats]
    its = torch.cat([r.iters for r in rows]).contiguous()
    sig = torch.cat([r.sigma for r in rows]).contiguous()
    bound = schedule.launch_bound([its], "sheet smoothing iterations")
    idx = (_lib.ctypes.c_void_p * sides)(*[_lib.ptr(st.sheet_indices) for st in stencils])
    wts = (_lib.ctypes.c_void_p * sides)(*[_lib.ptr(st.sheet_weights) for st in stencils])
    out = torch.empty((sides, nfw, st0.ny, st0.nx), dtype=out_dtype, device=st0.device)
    ws = _lib.workspace(_lib.lib().csg_type5_sheets_batch_workspace_bytes(
        n, topo.nnz, sides, nfw, m), st0.device)
    _lib.check(_lib.lib().csg_type5_sheets_batch_dev(
        _lib.ptr(g), _lib.ptr(sm), n, sides, _lib.ptr(topo.faces), _lib.ptr(topo.nbr_off),
        _lib.ptr(topo.nbr), topo.nnz, _lib.ptr(topo.inc_off), _lib.ptr(topo.inc), nfw,
        _lib.ptr(sig), _lib.ptr(its), int(bound), idx, wts, m, _lib.dtype_code(out_dtype),
        _lib.ptr(out), _lib.ptr(ws), ws.numel(), st0._stream()), "csg_type5_sheets_batch_dev")
    return out


def _minmax(values: torch.Tensor, floor: float = 0.0) -> torch.Tensor:
    """(x - min) / max(max - min, floor) on the device; a 2-D input per row."""

    x = values.to(torch.float64).contiguous()
    rows = 1 if x.ndim < 2 else int(x.shape[0])
    out = torch.empty_like(x)
    ws = _lib.workspace(_lib.lib().csg_minmax_workspace_bytes(rows), x.device)
    _lib.check(_lib.lib().csg_minmax_normalise(_lib.ptr(x), x.numel() // rows, rows,
                                               float(floor), 1, _lib.ptr(out), _lib.ptr(ws),
                                               ws.numel(), _lib.stream_handle(x.device)),
               "csg_minmax_normalise")
    return out
