"""In-process GPU stencil builder (replaces the CPU stencil helper).

The reference builds stencils with a native CPU helper linked against
CAT-Surface's libCAT (native/cat_surface_stencil_builder.c:342-464), run as a
subprocess 6 times per hemisphere.  This module builds the same stencil v2
content on the MI355X in-process:

- the 81,920-triangle icosahedral sphere ("tetrahedral sphere",
  create_tetrahedral_sphere in libCAT) at radius 0.975 x the mean input
  radius (c.f. :396-411);
- per resampled vertex: the closest triangle of the input sphere and the
  barycentric weights of the closest point, with indices into the surface
  (:107-135);
- the resampled sphere: input-sphere positions at those points;
- per map pixel: uv → point on the unit sphere (the inverse of the
  reference's own sphere→(u, v) map, surface_warp.py:526-535), closest
  triangle of the unit resampled sphere and weights (:137-170);
- the unit-sphere image of every input vertex (:172-185).

libCAT's source is not available, so its exact icosphere vertex order and
closest-point tie-breaking are not reproduced; parity of this builder with the
helper is unpinned (see the tests for the invariants it does satisfy).
"""

from __future__ import annotations

import math

import numpy as np
import torch

from . import _lib

from .locate import closest_triangles
from .surface_stencil import SurfaceStencil

_ICO_CACHE: dict[int, tuple[np.ndarray, np.ndarray]] = {}


def icosphere(subdivisions: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit icosphere by midpoint subdivision: 10·4^k + 2 vertices, 20·4^k faces."""

    if subdivisions in _ICO_CACHE:
        return _ICO_CACHE[subdivisions]
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
                      (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)],
                     dtype=np.float64)
    faces = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9),
                      (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2),
                      (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
                      (8, 6, 7), (9, 8, 1)], dtype=np.int64)
    verts /= np.linalg.norm(verts, axis=1, keepdims=True)
    for _ in range(subdivisions):
        nf = faces.shape[0]
        edges = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]])
        uniq, inverse = np.unique(np.sort(edges, axis=1), axis=0, return_inverse=True)
        inverse = inverse.reshape(-1)
        mids = verts[uniq[:, 0]] + verts[uniq[:, 1]]
        mids /= np.linalg.norm(mids, axis=1, keepdims=True)
        base = verts.shape[0]
        verts = np.concatenate([verts, mids])
        m01, m12, m20 = (base + inverse[k * nf:(k + 1) * nf] for k in range(3))
        faces = np.concatenate([np.stack([faces[:, 0], m01, m20], 1),
                                np.stack([faces[:, 1], m12, m01], 1),
                                np.stack([faces[:, 2], m20, m12], 1),
                                np.stack([m01, m12, m20], 1)])
    result = (np.ascontiguousarray(verts), np.ascontiguousarray(faces.astype(np.int32)))
    _ICO_CACHE[subdivisions] = result
    return result


def subdivisions_for(n_triangles: int) -> int:
    k = 0
    while 20 * 4 ** k < n_triangles:
        k += 1
    if 20 * 4 ** k != n_triangles:
        raise ValueError(f"n_triangles must be 20·4^k, got {n_triangles}")
    return k


def uv_points(nx: int, ny: int) -> np.ndarray:
    """Unit-sphere point of each pixel centre: u=(x+½)/nx → azimuth 2πu, v=(y+½)/ny → polar πv."""

    u = (np.arange(nx, dtype=np.float64) + 0.5) / nx
    v = (np.arange(ny, dtype=np.float64) + 0.5) / ny
    phi = (2.0 * np.pi * u)[None, :]
    theta = (np.pi * v)[:, None]
    pts = np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi),
                    np.broadcast_to(np.cos(theta), (ny, nx))], axis=-1)
    return np.ascontiguousarray(pts.reshape(-1, 3))


def build_stencil(sphere_vertices, sphere_faces, *, surface_vertices=None, normalise: bool = True,
                  nx: int = 512, ny: int = 256, n_triangles: int = 81920,
                  device: torch.device | None = None) -> SurfaceStencil:
    """Build a stencil v2 for a sphere mesh (and its surface); host copy for files."""

    d = build_device_stencil(sphere_vertices, sphere_faces, surface_vertices=surface_vertices,
                             normalise=normalise, nx=nx, ny=ny, n_triangles=n_triangles,
                             device=device)
    return SurfaceStencil(
        sphere_points=d.sphere_points64.cpu().numpy(),
        faces=d.topology.faces_np.copy(),
        surface_indices=d.surface_indices.cpu().numpy(),
        surface_weights=d.surface_weights.cpu().numpy(),
        sheet_indices=d.sheet_indices.cpu().numpy(),
        sheet_weights=d.sheet_weights.cpu().numpy(),
        nx=nx,
        ny=ny,
        source_points=d.source_points,
        unit_sphere_points=d.unit_sphere_points.cpu().numpy(),
    )


# Index grids of the two searches (cells per axis over [-1, 1]^3): the input
# sphere's faces for the resampled vertices, the resampled sphere's faces for
# the map pixels.  Any grid whose one-cell reach covers the queries gives the
# same stencil; finer ones scan fewer candidates but sort more pairs (measured
# at 61,442 vertices: 32/64 0.86 ms per build, 64/128 0.93 ms).
_GRID_SURFACE = 32
_GRID_SHEET = 64

_DEVICE_SPHERES: dict = {}
_DEVICE_UV: dict = {}


def _device_icosphere(subdivisions: int, dev: torch.device):
    """Icosphere vertices (float64), faces and topology on a device, built once."""

    from .surface_stencil import topology_for

    key = (dev.type, dev.index, subdivisions)
    entry = _DEVICE_SPHERES.get(key)
    if entry is None:
        ico_v, ico_f = icosphere(subdivisions)
        entry = (torch.as_tensor(ico_v, device=dev), torch.as_tensor(ico_f, device=dev),
                 topology_for(ico_f, int(ico_v.shape[0]), dev))
        _DEVICE_SPHERES[key] = entry
    return entry


def _device_uv(nx: int, ny: int, dev: torch.device) -> torch.Tensor:
    key = (dev.type, dev.index, nx, ny)
    uv = _DEVICE_UV.get(key)
    if uv is None:
        uv = torch.as_tensor(uv_points(nx, ny), device=dev)
        _DEVICE_UV[key] = uv
    return uv


def build_device_stencil(sphere_vertices, sphere_faces, *, surface_vertices=None,
                         normalise: bool = True, nx: int = 512, ny: int = 256,
                         n_triangles: int = 81920, device: torch.device | None = None,
                         deferred: bool = False):
    """Build the stencil directly as a SurfaceStencilDevice (no host round trip).

    ``sphere_vertices`` and ``sphere_faces`` may be host arrays or device
    tensors; the icosphere, its topology and the map's pixel directions are
    cached per device.  ``deferred=True`` returns (stencil, checks): nothing
    synchronises, and the stencil is exact iff ``locate.check_all(checks)``
    passes after the caller's next synchronisation (else build it again
    without ``deferred``).
    """

    from .dartel_grid import resolve_device
    from .surface_stencil import SurfaceStencilDevice

    dev = resolve_device("auto" if device is None else device)
    sv = torch.as_tensor(sphere_vertices, device=dev).to(torch.float32).contiguous()
    sf = torch.as_tensor(sphere_faces, device=dev).to(torch.int32).contiguous()
    if sv.ndim != 2 or sv.shape[1] != 3 or sf.ndim != 2 or sf.shape[1] != 3:
        raise ValueError("sphere vertices and faces must have shape [n, 3]")
    if surface_vertices is not None and tuple(np.shape(surface_vertices)) != tuple(sv.shape):
        raise ValueError("Surface and sphere vertex counts do not match")
    lib = _lib.lib()
    stream = _lib.stream_handle(dev)
    n = int(sv.shape[0])
    ico_v, ico_f, topo = _device_icosphere(subdivisions_for(n_triangles), dev)
    m = int(ico_v.shape[0])
    # float64 (normalised) sphere and the output sphere at 0.975 of its mean
    # radius around the bound sums (the helper's choice of centre)
    pts = torch.empty((n, 3), dtype=torch.float64, device=dev)
    out_sphere = torch.empty((m, 3), dtype=torch.float64, device=dev)
    params = torch.empty(8, dtype=torch.float64, device=dev)
    _lib.check(lib.csg_stencil_sphere(_lib.ptr(sv), n, int(bool(normalise)), _lib.ptr(ico_v), m,
                                      _lib.ptr(pts), _lib.ptr(out_sphere), _lib.ptr(params),
                                      stream), "csg_stencil_sphere")
    # surface stencil: closest input-sphere triangle per resampled vertex
    # (the result is the global closest triangle, first by face order on ties,
    # whatever the grid: the grid only sets how many candidates a query scans;
    # the output sphere lies 2.5 % inside the input sphere, within one cell)
    first = closest_triangles(out_sphere, pts, sf, grid_size=_GRID_SURFACE, margin=1,
                              deferred=deferred)
    surf_idx, surf_w = first[1], first[2]
    resampled = torch.empty((m, 3), dtype=torch.float64, device=dev)
    unit_res = torch.empty((m, 3), dtype=torch.float64, device=dev)
    _lib.check(lib.csg_resample_unit(_lib.ptr(pts), _lib.ptr(surf_idx), _lib.ptr(surf_w), m,
                                     _lib.ptr(resampled), _lib.ptr(unit_res), stream),
               "csg_resample_unit")
    # sheet stencil on the unit resampled sphere: the pixel directions lie on
    # that mesh (within its chords' sagitta), so the block search on an index
    # without margin decides almost every query in its own cell
    second = closest_triangles(_device_uv(nx, ny, dev), unit_res, ico_f,
                               grid_size=_GRID_SHEET, margin=1, deferred=deferred, block=True)
    sheet_idx, sheet_w = second[1], second[2]
    unit_pts = torch.empty((n, 3), dtype=torch.float32, device=dev)
    _lib.check(lib.csg_unit_points64(_lib.ptr(pts), n, _lib.ptr(unit_pts), stream),
               "csg_unit_points64")
    stencil = SurfaceStencilDevice(
        sphere_points=resampled.to(torch.float32),
        faces=topo.faces,
        surface_indices=surf_idx.contiguous(),
        surface_weights=surf_w.contiguous(),
        sheet_indices=sheet_idx.reshape(ny, nx, 3).contiguous(),
        sheet_weights=sheet_w.reshape(ny, nx, 3).contiguous(),
        nx=nx,
        ny=ny,
        source_points=int(sv.shape[0]),
        unit_sphere_points=unit_pts,
        topology=topo,
        sphere_points64=resampled,
    )
    return (stencil, [first[3], second[3]]) if deferred else stencil
