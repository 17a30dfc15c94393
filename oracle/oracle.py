"""ctypes front end of the CPU oracle (test infrastructure only).

The oracle restates the reference's CPU arithmetic in plain C (csg_oracle.c,
csg_oracle_surface.c) and in small numpy helpers here.  Only tests/, the smoke
entry point and bench.py's CPU-baseline leg may import this module; the product
package never does.

Build: ``make -C oracle`` (``__graft_entry__.build()`` runs it).
"""

from __future__ import annotations

import ctypes
import hashlib
import math
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB: ctypes.CDLL | None = None

_D = ctypes.POINTER(ctypes.c_double)
_F = ctypes.POINTER(ctypes.c_float)
_I = ctypes.POINTER(ctypes.c_int)
_L = ctypes.POINTER(ctypes.c_long)


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        path = _HERE / "liboracle.so"
        if not path.is_file():
            raise RuntimeError(f"oracle library is not built: run `make -C {_HERE}`")
        _LIB = ctypes.CDLL(str(path))
        _LIB.oracle_initialise_objfun.restype = ctypes.c_double
    return _LIB


def _d(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def _f(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_F)


def _i(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_I)


def _c64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _c32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _ci(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _shape(field) -> tuple[int, int]:
    return int(field.shape[-1]), int(field.shape[-2])


# ---------------------------------------------------------------- DARTEL grid


def bound(ix, iy, nx: int, ny: int):
    ix = np.ascontiguousarray(np.asarray(ix, dtype=np.int64))
    iy = np.ascontiguousarray(np.asarray(iy, dtype=np.int64))
    ox = np.empty(ix.shape, np.int32)
    oy = np.empty(iy.shape, np.int32)
    lib().oracle_bound(ix.ctypes.data_as(_L), iy.ctypes.data_as(_L), ctypes.c_int(ix.size),
                       nx, ny, _i(ox), _i(oy))
    return ox, oy


def make_identity_map(nx: int, ny: int) -> np.ndarray:
    x = np.arange(1, nx + 1, dtype=np.float64)
    y = np.arange(1, ny + 1, dtype=np.float64)
    return np.stack([np.broadcast_to(x[None, :], (ny, nx)), np.broadcast_to(y[:, None], (ny, nx))]).copy()


def sample_field(field, x, y):
    field = _c64(field)
    x = _c64(x)
    y = _c64(y)
    out = np.empty(x.shape, np.float64)
    nx, ny = _shape(field)
    lib().oracle_sample_field(_d(field), nx, ny, _d(x), _d(y), ctypes.c_int(x.size), _d(out))
    return out


def composition(a, b):
    a = _c64(a)
    b = _c64(b)
    nx, ny = _shape(a)
    out = np.empty_like(a)
    lib().oracle_composition(_d(a), _d(b), nx, ny, _d(out))
    return out


def composition_jacobian(a, ja, b, jb):
    a, ja, b, jb = (_c64(t) for t in (a, ja, b, jb))
    nx, ny = _shape(a)
    m = np.empty_like(a)
    j = np.empty_like(ja)
    lib().oracle_composition_jacobian(_d(a), _d(ja), _d(b), _d(jb), nx, ny, _d(m), _d(j))
    return m, j


def jacobian_of_displacement(v, scale: float = 1.0):
    v = _c64(v)
    nx, ny = _shape(v)
    out = np.empty((4, ny, nx), np.float64)
    lib().oracle_jacobian_of_displacement(_d(v), ctypes.c_double(scale), nx, ny, _d(out))
    return out


def smalldef_jac(v, scale: float):
    v = _c64(v)
    nx, ny = _shape(v)
    m = np.empty((2, ny, nx), np.float64)
    j = np.empty((4, ny, nx), np.float64)
    lib().oracle_smalldef_jac(_d(v), ctypes.c_double(scale), nx, ny, _d(m), _d(j))
    return m, j


def jac_div_smalldef(jac, v, scale: float):
    jac = _c64(jac)
    v = _c64(v)
    nx, ny = _shape(v)
    out = np.empty((4, ny, nx), np.float64)
    lib().oracle_jac_div_smalldef(_d(jac), _d(v), ctypes.c_double(scale), nx, ny, _d(out))
    return out


def expdef(v, k: int, return_jacobian: bool = False):
    v = _c64(v)
    nx, ny = _shape(v)
    m = np.empty((2, ny, nx), np.float64)
    j = np.empty((4, ny, nx), np.float64)
    lib().oracle_expdef(_d(v), k, int(return_jacobian), nx, ny, _d(m), _d(j))
    return (m, j) if return_jacobian else m


def initialise_objfun(src, tgt, phi, jac, dist=None):
    src, tgt, phi, jac = (_c64(t) for t in (src, tgt, phi, jac))
    nx, ny = _shape(src)
    rhs = np.empty((2, ny, nx), np.float64)
    hes = np.empty((3, ny, nx), np.float64)
    dptr = _d(_c64(dist)) if dist is not None else None
    if dist is not None:
        dist = _c64(dist)
        dptr = _d(dist)
    obj = lib().oracle_initialise_objfun(_d(src), _d(tgt), _d(phi), _d(jac), dptr, nx, ny,
                                         _d(rhs), _d(hes))
    return obj, rhs, hes


def squaring_update(rhs, hes, phi, jac, k: int, save_transformation: bool = False):
    rhs, hes, phi, jac = (_c64(t) for t in (rhs, hes, phi, jac))
    nx, ny = _shape(phi)
    b = np.empty_like(rhs)
    a = np.empty_like(hes)
    m = np.empty_like(phi)
    j = np.empty_like(jac)
    s = np.empty_like(phi)
    lib().oracle_squaring_update(_d(rhs), _d(hes), _d(phi), _d(jac), k, int(save_transformation),
                                 nx, ny, _d(b), _d(a), _d(m), _d(j), _d(s))
    return b, a, m, j, s


def regularization_operator(v, params, rtype: int = 1):
    v = _c64(v)
    p = _c64(params)
    nx, ny = _shape(v)
    out = np.empty_like(v)
    lib().oracle_regularization(_d(v), _d(p), rtype, nx, ny, _d(out))
    return out


def apply_membrane_system(hes, u, params):
    hes, u, p = _c64(hes), _c64(u), _c64(params)
    nx, ny = _shape(u)
    out = np.empty_like(u)
    lib().oracle_apply_membrane(_d(hes), _d(u), _d(p), nx, ny, _d(out))
    return out


def relax_membrane(hes, rhs, params, nit: int, initial=None):
    hes, rhs, p = _c64(hes), _c64(rhs), _c64(params)
    nx, ny = _shape(rhs)
    u = np.zeros_like(rhs) if initial is None else _c64(initial).copy()
    lib().oracle_relax(_d(hes), _d(rhs), _d(p), nit, nx, ny, _d(u))
    return u


def resize_field(field, tnx: int, tny: int):
    f = _c64(field)
    if f.ndim == 2:
        f = f[None]
    c = f.shape[0]
    snx, sny = _shape(f)
    out = np.empty((c, tny, tnx), np.float64)
    lib().oracle_resize(_d(f), c, snx, sny, tnx, tny, _d(out))
    return out


def fmg2_membrane(hes, rhs, params, cycles: int = 3, nit: int = 3, initial=None,
                  assume_initial_nonzero: bool = False):
    hes, rhs, p = _c64(hes), _c64(rhs), _c64(params)
    nx, ny = _shape(rhs)
    out = np.empty_like(rhs)
    init = None
    if initial is not None:
        init_arr = _c64(initial)
        init = _d(init_arr)
    lib().oracle_fmg2(_d(hes), _d(rhs), _d(p), cycles, nit, init, int(assume_initial_nonzero),
                      nx, ny, _d(out))
    return out


def dartel_step(src, tgt, v, k: int, params=(1.0, 1.0, 0.125, 0.0, 0.0), lmreg: float = 1e-3,
                cycles: int = 3, nit: int = 3, code: int = 1, distortion=None,
                assume_initial_nonzero: bool = False):
    src, tgt, v, p = _c64(src), _c64(tgt), _c64(v), _c64(params)
    nx, ny = _shape(v)
    out = np.empty_like(v)
    metrics = np.empty(3, np.float64)
    dptr = None
    if distortion is not None:
        dist = _c64(distortion).reshape(ny, nx)
        dptr = _d(dist)
    lib().oracle_dartel_step(_d(src), _d(tgt), _d(v), k, _d(p), ctypes.c_double(lmreg), cycles,
                             nit, code, dptr, int(assume_initial_nonzero), nx, ny, _d(out),
                             _d(metrics))
    return out, metrics


def default_dartel_parameters(loop: int = 6) -> np.ndarray:
    out = np.empty((loop, 5), np.float64)
    lib().oracle_default_params(loop, _d(out))
    return out


def solve_dartel_maps(src, tgt, loop: int = 6, lmreg: float = 1e-3, cycles: int = 3,
                      nit: int = 3, its: int = 3, code: int = 1,
                      assume_initial_nonzero: bool = False):
    src, tgt = _c64(src), _c64(tgt)
    steps = src.shape[0]
    nx, ny = _shape(src)
    flow = np.empty((2, ny, nx), np.float64)
    metrics = np.empty((steps, loop * its, 3), np.float64)
    lib().oracle_solve_dartel_maps(_d(src), _d(tgt), steps, loop, ctypes.c_double(lmreg), cycles,
                                   nit, its, code, int(assume_initial_nonzero), nx, ny, _d(flow),
                                   _d(metrics))
    return flow, metrics


# ---------------------------------------------------------------- warp


def wrap_displacement(values):
    v = _c64(values)
    out = np.empty_like(v)
    lib().oracle_wrap_displacement(_d(v), ctypes.c_int(v.size), _d(out))
    return out


def apply_flow_to_sphere(unit_points, flow, inverse: bool = True):
    pts = _c32(unit_points)
    flow = _c64(flow)
    nx, ny = _shape(flow)
    out = np.empty((pts.shape[0], 3), np.float64)
    lib().oracle_apply_flow(_f(pts), ctypes.c_int(pts.shape[0]), _d(flow), nx, ny,
                            int(inverse), _d(out))
    return out


def normalise_points_f32(points) -> np.ndarray:
    """Per-vertex float32 normalisation used when no unit points are cached.

    src/cat_surface_gpu/surface_warp.py:518-522
    """
    p = _c32(points)
    out = np.empty_like(p)
    lib().oracle_normalise_f32(_f(p), ctypes.c_int(p.shape[0]), _f(out))
    return out


def average_xz_surfaces(x_surface, z_surface):
    xs, zs = _c64(x_surface), _c64(z_surface)
    out = np.empty(xs.shape, np.float32)
    lib().oracle_average_xz(_d(xs), _d(zs), ctypes.c_int(xs.shape[0]), _f(out))
    return out


def rotation_matrix_from_angles(angles) -> np.ndarray:
    """Rz (Ry Rx) assembled column-major then read row-major.

    src/cat_surface_gpu/gpu_pipeline.py:308-370
    """
    alpha, beta, gamma = (float(a) for a in np.asarray(angles, dtype=np.float64).reshape(3))
    ca, sa, cb, sb, cg, sg = (math.cos(alpha), math.sin(alpha), math.cos(beta), math.sin(beta),
                              math.cos(gamma), math.sin(gamma))
    rx = [1.0, 0.0, 0.0, 0.0, ca, sa, 0.0, -sa, ca]
    ry = [cb, 0.0, sb, 0.0, 1.0, 0.0, -sb, 0.0, cb]
    rz = [cg, sg, 0.0, -sg, cg, 0.0, 0.0, 0.0, 1.0]
    first = [0.0] * 9
    result = [0.0] * 9
    for row in range(3):
        for col in range(3):
            first[row + 3 * col] = sum(ry[row + 3 * i] * rx[i + 3 * col] for i in range(3))
    for row in range(3):
        for col in range(3):
            result[row + 3 * col] = sum(rz[row + 3 * i] * first[i + 3 * col] for i in range(3))
    return np.asarray(result, dtype=np.float64).reshape(3, 3)


def normalise_sphere(vertices) -> np.ndarray:
    """Sequential FP64 centroid, then per-vertex float32 normalisation.

    src/cat_surface_gpu/gpu_pipeline.py:274-305
    """
    v = _c32(vertices)
    # np.cumsum accumulates strictly in order, as the reference's Python loop does
    centre = np.cumsum(v.astype(np.float64), axis=0)[-1]
    centred = (v.astype(np.float64) - centre / v.shape[0]).astype(np.float32)
    c64 = centred.astype(np.float64)
    length = np.sqrt((c64 * c64).sum(1))
    if not np.all(np.isfinite(length)) or np.any(length == 0.0):
        raise ValueError("Sphere contains a vertex that cannot be normalized")
    return (c64 * (1.0 / length)[:, None]).astype(np.float32)


def rotate_vertices(vertices, angles) -> np.ndarray:
    m = rotation_matrix_from_angles(angles)
    return (np.asarray(vertices, dtype=np.float64) @ m.T).astype(np.float32)


# ---------------------------------------------------------------- surface side


_CSR_CACHE: dict = {}


def neighbour_csr(faces, n_points: int):
    """Ascending neighbour lists per vertex (surface_stencil.py:36-64)."""

    f = np.ascontiguousarray(np.asarray(faces, dtype=np.int64))
    key = (hashlib.sha1(f.tobytes()).hexdigest(), f.shape, int(n_points))
    hit = _CSR_CACHE.get(key)
    if hit is not None:
        return hit
    src = np.concatenate([f[:, 0], f[:, 1], f[:, 1], f[:, 2], f[:, 2], f[:, 0]])
    dst = np.concatenate([f[:, 1], f[:, 0], f[:, 2], f[:, 1], f[:, 0], f[:, 2]])
    # unique directed edges sorted by (source, target), as one integer key each
    pairs = np.unique(src * np.int64(n_points) + dst)
    counts = np.bincount(pairs // n_points, minlength=n_points)
    off = np.zeros(n_points + 1, np.int32)
    off[1:] = np.cumsum(counts)
    out = (off, np.ascontiguousarray((pairs % n_points).astype(np.int32)))
    if len(_CSR_CACHE) > 16:
        _CSR_CACHE.clear()
    _CSR_CACHE[key] = out
    return out


def resample_vertices_f32(x, idx, w):
    x, idx, w = _c32(x), _ci(idx), _c64(w)
    out = np.empty((idx.shape[0], 3), np.float32)
    lib().oracle_resample_f32(_f(x), _i(idx), _d(w), ctypes.c_int(idx.shape[0]), _f(out))
    return out


def edge_mean(vertices, faces) -> float:
    lib().oracle_edge_mean.restype = ctypes.c_double
    v, f = _c32(vertices), _ci(faces)
    return float(lib().oracle_edge_mean(_f(v), _i(f), ctypes.c_int(f.shape[0])))


def heat_parameters(vertices, faces, fwhm: float):
    em = edge_mean(vertices, faces)
    it = ctypes.c_int()
    sg = ctypes.c_double()
    lib().oracle_heat_parameters(ctypes.c_double(em), ctypes.c_double(fwhm), ctypes.byref(it),
                                 ctypes.byref(sg))
    return it.value, sg.value


def smooth_geometry(vertices, faces, fwhm: float):
    v = _c32(vertices)
    off, nbr = neighbour_csr(faces, v.shape[0])
    iters, _ = heat_parameters(v, faces, fwhm)
    out = np.empty_like(v)
    lib().oracle_smooth_geometry(_f(v), _i(off), _i(nbr), ctypes.c_int(v.shape[0]), iters,
                                 _f(out))
    return out


def smooth_values(values, vertices, faces, fwhm: float, squared_norm: bool = False):
    vals = _c64(values)
    geom32 = _c32(vertices)
    geom = _c64(geom32)
    off, nbr = neighbour_csr(faces, vals.shape[0])
    iters, sigma = heat_parameters(geom32, faces, fwhm)
    w = np.empty(nbr.shape[0], np.float64)
    lib().oracle_heat_weights(_d(geom), _i(off), _i(nbr), ctypes.c_int(vals.shape[0]),
                              ctypes.c_double(sigma), int(squared_norm), _d(w))
    out = np.empty_like(vals)
    lib().oracle_smooth_values(_d(vals), _i(off), _i(nbr), _d(w), ctypes.c_int(vals.shape[0]),
                               iters, _d(out))
    return out


def vertex_normals(vertices, faces):
    v, f = _c32(vertices), _ci(faces)
    out = np.empty_like(v)
    lib().oracle_vertex_normals(_f(v), _i(f), ctypes.c_int(f.shape[0]), ctypes.c_int(v.shape[0]),
                                _f(out))
    return out


def sheet_gather(values, idx, w):
    vals, idx, w = _c64(values), _ci(idx), _c64(w)
    m = idx.size // 3
    out = np.empty(m, np.float64)
    lib().oracle_sheet_gather(_d(vals), _i(idx), _d(w), ctypes.c_int(m), _d(out))
    return out


def minmax_normalise(x):
    x = _c64(x).copy()
    lib().oracle_minmax_normalise(_d(x), ctypes.c_int(x.size))
    return x


def type5_base(geometry, faces):
    g = _c32(geometry)
    sm = smooth_geometry(g, faces, 50.0)
    nr = vertex_normals(sm, faces)
    out = np.empty(g.shape[0], np.float64)
    lib().oracle_type5_base(_f(g), _f(sm), _f(nr), ctypes.c_int(g.shape[0]), _d(out))
    return out


def curvature_type2(geometry, faces):
    """Type-2 curvature before smoothing (surface_stencil.py:552-613)."""

    g = _c32(geometry)
    nr = vertex_normals(g, faces)
    off, nbr = neighbour_csr(faces, g.shape[0])
    out = np.empty(g.shape[0], np.float64)
    lib().oracle_curvature_type2(_f(g), _f(nr), _i(off), _i(nbr), ctypes.c_int(g.shape[0]),
                                 _d(out))
    return out


def curvature_to_sheet(geometry, faces, sheet_idx, sheet_w, fwhm: float, curvtype: int = 5):
    """curvature_type5 (surface_stencil.py:507-519) or curvature_type2 (:552-622)
    for one fwhm, smoothed with |d|^2 weights, normalised, gathered to the sheet
    and normalised again (:624-644)."""

    g = _c32(geometry)
    if curvtype == 5:
        raw = type5_base(g, faces)
    elif curvtype == 2:
        raw = curvature_type2(g, faces)
    else:
        raise NotImplementedError("only upstream curvtype 2 and 5 exist")
    vals = minmax_normalise(smooth_values(raw, g, faces, fwhm, squared_norm=False))
    ny, nx = sheet_idx.shape[:2]
    sheet = sheet_gather(vals, sheet_idx.reshape(-1, 3), sheet_w.reshape(-1, 3))
    return minmax_normalise(sheet).reshape(ny, nx)


def curvature_type5_to_sheet_many(geometry, faces, sheet_idx, sheet_w, fwhms):
    """surface_stencil.py:521-550"""

    g = _c32(geometry)
    base = type5_base(g, faces)
    ny, nx = sheet_idx.shape[:2]
    maps = []
    for fwhm in fwhms:
        vals = smooth_values(base, g, faces, fwhm, squared_norm=True)
        vals = minmax_normalise(vals)
        sheet = sheet_gather(vals, sheet_idx.reshape(-1, 3), sheet_w.reshape(-1, 3))
        maps.append(minmax_normalise(sheet).reshape(ny, nx))
    return maps


# ---------------------------------------------------------------- rotation


def official_seed_grid_angles() -> np.ndarray:
    """surface_rotation.py:14-23"""

    seed0 = (-0.6, -0.3, 0.0, 0.3, 0.6)
    seed12 = (-0.3, 0.0, 0.3)
    values = [(0.0, 0.0, 0.0)] + [(a, b, c) for a in seed0 for b in seed12 for c in seed12]
    return np.asarray(values, dtype=np.float64)


def grid_index(points, faces, grid_size: int = 128, margin: int = 1):
    """CSR of triangles per cell of a grid on [-1, 1]^3 (surface_rotation.py:57-120)."""

    points = np.asarray(points, dtype=np.float64)
    faces = np.asarray(faces, dtype=np.int64)
    cell = 2.0 / float(grid_size)
    tri = points[faces]
    lo = np.floor((tri.min(axis=1) + 1.0) / cell).astype(np.int64)
    hi = np.floor((tri.max(axis=1) + 1.0) / cell).astype(np.int64)
    lo = np.clip(lo - margin, 0, grid_size - 1)
    hi = np.clip(hi + margin, 0, grid_size - 1)
    cells, owners = [], []
    for f in range(faces.shape[0]):
        zs = np.arange(lo[f, 2], hi[f, 2] + 1)
        ys = np.arange(lo[f, 1], hi[f, 1] + 1)
        xs = np.arange(lo[f, 0], hi[f, 0] + 1)
        ids = (xs[None, None, :] + grid_size * (ys[None, :, None] + grid_size * zs[:, None, None]))
        ids = ids.reshape(-1)
        cells.append(ids)
        owners.append(np.full(ids.size, f, np.int64))
    cells = np.concatenate(cells)
    owners = np.concatenate(owners)
    order = np.lexsort((owners, cells))
    counts = np.bincount(cells, minlength=grid_size ** 3)
    off = np.zeros(grid_size ** 3 + 1, np.int32)
    off[1:] = np.cumsum(counts)
    return off, owners[order].astype(np.int32)


def rotation_costs(src_pts, src_vals, tgt_vals, points, faces, angles, grid_size=128, margin=1,
                   index=None, return_location=False):
    src_pts, src_vals, tgt_vals = _c32(src_pts), _c64(src_vals), _c64(tgt_vals)
    faces = _ci(faces)
    tri = _c32(np.asarray(points, dtype=np.float64)[faces])
    off, cf = index if index is not None else grid_index(points, faces, grid_size, margin)
    off, cf = _ci(off), _ci(cf)
    angles = _c64(np.asarray(angles, dtype=np.float64).reshape(-1, 3))
    na, ns = angles.shape[0], src_pts.shape[0]
    costs = np.empty(na, np.float64)
    lf = np.empty((na, ns), np.int32)
    lb = np.empty((na, ns, 3), np.float32)
    rc = lib().oracle_rotation_costs(_f(src_pts), _d(src_vals), ctypes.c_int(ns), _d(tgt_vals),
                                     _i(faces), _f(tri), _i(off), _i(cf), grid_size, _d(angles),
                                     ctypes.c_int(na), _d(costs), _i(lf), _f(lb))
    if rc != 0:
        raise RuntimeError("Rotation is not covered by target sphere candidate index")
    if return_location:
        return costs, lf, lb
    return costs


def rotation_matrix_f32(angles) -> np.ndarray:
    a = _c64(angles)
    m = np.empty((3, 3), np.float32)
    lib().oracle_rotation_matrix_f32(_d(a), _f(m))
    return m


def nelder_mead(cost, seed, max_iter=500, tol=1e-4, simplex_step=0.1):
    """CAT-ordered Nelder-Mead (surface_rotation.py:594-699); cost(angles[k,3]) -> [k]."""

    seed = np.asarray(seed, dtype=np.float64).reshape(3)
    simplex = np.stack([seed, seed + (simplex_step, 0, 0), seed + (0, simplex_step, 0),
                        seed + (0, 0, simplex_step)])
    f = np.array(cost(simplex), dtype=np.float64)
    iteration = 0
    for iteration in range(max_iter):
        hi, lo, sh = 0, 0, 1
        for i in range(4):
            if f[i] > f[hi]:
                sh, hi = hi, i
            elif f[i] > f[sh] and i != hi:
                sh = i
            if f[i] < f[lo]:
                lo = i
        centroid = np.zeros(3)
        for i in range(4):
            if i != hi:
                centroid += simplex[i]
        centroid /= 3.0
        refl = centroid + (centroid - simplex[hi])
        fr = cost(refl[None])[0]
        if fr < f[lo]:
            exp = centroid + 2.0 * (refl - centroid)
            fe = cost(exp[None])[0]
            if fe < fr:
                simplex[hi], f[hi] = exp, fe
            else:
                simplex[hi], f[hi] = refl, fr
        elif fr < f[sh]:
            simplex[hi], f[hi] = refl, fr
        else:
            con = centroid + 0.5 * (simplex[hi] - centroid)
            fc = cost(con[None])[0]
            if fc < f[hi]:
                simplex[hi], f[hi] = con, fc
            else:
                for i in range(4):
                    if i != lo:
                        simplex[i] = simplex[lo] + 0.5 * (simplex[i] - simplex[lo])
                        f[i] = cost(simplex[i][None])[0]
        if np.max(np.abs(f - f[lo])) < tol:
            break
    return simplex[lo].copy(), float(f[lo]), iteration + 1


def closest_triangles(verts, faces, queries):
    v, f, q = _c64(verts), _ci(faces), _c64(queries)
    fo = np.empty(q.shape[0], np.int32)
    w = np.empty((q.shape[0], 3), np.float64)
    lib().oracle_closest_triangles(_d(v), _i(f), ctypes.c_int(f.shape[0]), _d(q),
                                   ctypes.c_int(q.shape[0]), _i(fo), _d(w))
    return fo, w
