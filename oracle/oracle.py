"""ctypes front end of the CPU oracle (test infrastructure only).

The oracle restates the reference's CPU arithmetic in plain C (csg_oracle.c,
csg_oracle_surface.c) and in small numpy helpers here.  Only tests/, the smoke
entry point and bench.py's CPU-baseline leg may import this module; the product
package never does.

Build: ``make -C oracle`` (``__graft_entry__.build()`` runs it).
"""

from __future__ import annotations

import ctypes
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
