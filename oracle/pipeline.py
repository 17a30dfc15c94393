"""End-to-end CAT_SurfWarp on the CPU oracle (test infrastructure only).

Composes the oracle's pinned stages into the whole registration the reference's
``run_cat_surface_gpu_pipeline`` performs (src/cat_surface_gpu/gpu_pipeline.py:
468-1326): stencils of the input spheres, rotation features, the 46-seed scan and
Nelder-Mead, the rotated stencils, the DARTEL runs on type-5 curvature sheets,
the inverse warps, the averaged pole-frame solve and its combination.  Only
tests/, the smoke entry point and bench.py's CPU-baseline leg import this
module; the product package never does.

The reference builds its stencils with a helper linked against CAT-Surface's
libCAT, whose source is absent; the stencil restated here follows the same
construction as the GPU builder (icosphere of 81,920 triangles at 0.975 of the
mean radius around the bound sums, exact closest triangles with barycentric
weights, equirectangular pixel directions), so that the chain checks the
registration path, not a different stencil definition.
"""

from __future__ import annotations

import ctypes
import math
import resource
import time
from dataclasses import dataclass, field

import numpy as np

from oracle import oracle as O

_EPS64 = 2.220446049250313e-16
POLE = np.asarray((0.0, math.pi / 2.0, 0.0), dtype=np.float64)


def _keep_freed_memory() -> None:
    """The oracle's DARTEL operators allocate and free map-sized buffers at every
    call; keeping them in the heap (no mmap/munmap and page faults per call)
    makes a full solve several times faster.  Values are unaffected."""

    try:
        libc = ctypes.CDLL("libc.so.6")
        libc.mallopt(-3, 1 << 30)   # M_MMAP_THRESHOLD
        libc.mallopt(-1, (1 << 31) - 1)   # M_TRIM_THRESHOLD
        libc.mallopt(-2, 1 << 28)   # M_TOP_PAD
    except OSError:
        pass


_keep_freed_memory()


# ---------------------------------------------------------------- stencil


def icosphere(subdivisions: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit icosphere by midpoint subdivision (10·4^k + 2 vertices, 20·4^k faces);
    new vertices are numbered by sorted edge, children of face f are f, f + F,
    f + 2F (corner 0, 1, 2) and f + 3F (centre)."""

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
        edges = np.sort(np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]]),
                        axis=1)
        uniq, inverse = np.unique(edges, axis=0, return_inverse=True)
        inverse = inverse.reshape(-1)
        mids = verts[uniq[:, 0]] + verts[uniq[:, 1]]
        mids /= np.linalg.norm(mids, axis=1, keepdims=True)
        base = verts.shape[0]
        verts = np.concatenate([verts, mids])
        a, b, c = (base + inverse[k * nf:(k + 1) * nf] for k in range(3))
        faces = np.concatenate([np.stack([faces[:, 0], a, c], 1),
                                np.stack([faces[:, 1], b, a], 1),
                                np.stack([faces[:, 2], c, b], 1),
                                np.stack([a, b, c], 1)])
    return np.ascontiguousarray(verts), np.ascontiguousarray(faces.astype(np.int32))


def uv_points(nx: int, ny: int) -> np.ndarray:
    """Unit direction of each pixel centre: azimuth 2π(x+½)/nx, polar angle π(y+½)/ny
    (the inverse of the sphere→map coordinates of surface_warp.py:526-535)."""

    az = 2.0 * np.pi * (np.arange(nx, dtype=np.float64) + 0.5) / nx
    po = np.pi * (np.arange(ny, dtype=np.float64) + 0.5) / ny
    s, c = np.sin(po)[:, None], np.cos(po)[:, None]
    pts = np.stack([s * np.cos(az)[None, :], s * np.sin(az)[None, :],
                    np.broadcast_to(c, (ny, nx))], axis=-1)
    return np.ascontiguousarray(pts.reshape(-1, 3))


def closest_triangles(verts, faces, queries) -> tuple[np.ndarray, np.ndarray]:
    """Exact closest triangle of each query (first minimum by face index) and the
    barycentric weights of the closest point, through a shell-searched grid."""

    v, f, q = O._c64(verts), O._ci(faces), O._c64(queries)
    face = np.empty(q.shape[0], np.int32)
    w = np.empty((q.shape[0], 3), np.float64)
    rc = O.lib().oracle_closest_triangles_grid(O._d(v), ctypes.c_int(v.shape[0]), O._i(f),
                                               ctypes.c_int(f.shape[0]), O._d(q),
                                               ctypes.c_int(q.shape[0]), O._i(face), O._d(w))
    if rc != 0:
        raise RuntimeError(f"oracle closest-triangle search failed ({rc})")
    return face, w


@dataclass
class Stencil:
    sphere64: np.ndarray      # [R, 3] resampled sphere (input-sphere positions)
    faces: np.ndarray         # [F, 3] icosphere faces
    surf_idx: np.ndarray      # [R, 3] input vertex indices
    surf_w: np.ndarray        # [R, 3] barycentric weights
    sheet_idx: np.ndarray     # [ny, nx, 3] resampled-sphere vertex indices
    sheet_w: np.ndarray       # [ny, nx, 3]
    unit_points: np.ndarray   # [N, 3] float32 unit image of every input vertex


def build_stencil(sphere_vertices, sphere_faces, *, normalise: bool, nx: int = 512,
                  ny: int = 256, n_triangles: int = 81920) -> Stencil:
    """Stencil v2 content of a sphere mesh (the stencil helper's construction,
    native/cat_surface_stencil_builder.c:107-185, :342-464)."""

    k = 0
    while 20 * 4 ** k < n_triangles:
        k += 1
    if 20 * 4 ** k != n_triangles:
        raise ValueError("n_triangles must be 20·4^k")
    pts = np.asarray(sphere_vertices, dtype=np.float32).astype(np.float64)
    faces = O._ci(sphere_faces)
    if normalise:
        pts = pts - pts.sum(0) / pts.shape[0]
        pts = pts / np.sqrt((pts * pts).sum(1))[:, None]
    radius = 0.975 * np.sqrt((pts * pts).sum(1)).mean()
    ico_v, ico_f = icosphere(k)
    out_sphere = (pts.min(0) + pts.max(0))[None, :] + radius * ico_v
    face, w = closest_triangles(pts, faces, out_sphere)
    surf_idx = np.ascontiguousarray(faces[face])
    res = ((pts[surf_idx[:, 0]] * w[:, 0:1] + pts[surf_idx[:, 1]] * w[:, 1:2])
           + pts[surf_idx[:, 2]] * w[:, 2:3])
    unit_res = res / np.sqrt((res * res).sum(1))[:, None]
    face2, w2 = closest_triangles(unit_res, ico_f, uv_points(nx, ny))
    unit = (pts / np.sqrt((pts * pts).sum(1))[:, None]).astype(np.float32)
    return Stencil(sphere64=res, faces=ico_f, surf_idx=surf_idx, surf_w=w,
                   sheet_idx=np.ascontiguousarray(ico_f[face2].reshape(ny, nx, 3)),
                   sheet_w=np.ascontiguousarray(w2.reshape(ny, nx, 3)),
                   unit_points=np.ascontiguousarray(unit))


# ---------------------------------------------------------------- rotation features


def greedy_colours(faces, n_points: int) -> np.ndarray:
    """Greedy colouring in vertex order (surface_stencil.py:67-90)."""

    off, nbr = O.neighbour_csr(faces, n_points)
    colours = np.empty(n_points, np.int32)
    count = O.lib().oracle_greedy_colours(O._i(off), O._i(nbr), ctypes.c_int(n_points),
                                          O._i(colours))
    if count < 1:
        raise RuntimeError("colouring failed")
    return colours


@dataclass
class DepthResult:
    values: np.ndarray
    iterations: int
    relative_residual: float
    normals: np.ndarray
    areas: np.ndarray


def depth_potential(geometry, faces, *, solver: str = "colored-sor", alpha: float = 1e-3,
                    max_iter: int = 1000, relaxation: float = 1.90, tolerance: float = 1e-9,
                    colours=None) -> DepthResult:
    """Depth potential of a smoothed surface (rotation_feature.py:193-317)."""

    g = O._c32(geometry)
    f = O._ci(faces)
    n = g.shape[0]
    if colours is None:
        colours = greedy_colours(f, n)
    colours = O._ci(colours)
    x = np.empty(n, np.float64)
    normals = np.empty((n, 3), np.float64)
    areas = np.empty(n, np.float64)
    rel = ctypes.c_double()
    lib = O.lib()
    lib.oracle_depth_potential.restype = ctypes.c_int
    its = lib.oracle_depth_potential(
        O._f(g), O._i(f), ctypes.c_int(f.shape[0]), ctypes.c_int(n), ctypes.c_double(alpha),
        ctypes.c_int(0 if solver == "colored-sor" else 1), ctypes.c_int(max_iter),
        ctypes.c_double(relaxation), ctypes.c_double(tolerance), O._i(colours),
        ctypes.c_int(int(colours.max()) + 1), O._d(x), O._d(normals), O._d(areas),
        ctypes.byref(rel))
    return DepthResult(x, int(its), float(rel.value), normals, areas)


def rotation_feature(stencil: Stencil, surface_vertices, heat_fwhm: float, *,
                     solver: str = "colored-sor", max_iter: int = 1000) -> np.ndarray:
    """The reference's 'cuda' rotation feature (rotation_feature.py:320-399):
    resample, heat-smooth the geometry, depth potential, fwhm-50 value smoothing,
    min/max normalisation."""

    mapped = O.resample_vertices_f32(surface_vertices, stencil.surf_idx, stencil.surf_w)
    smoothed = O.smooth_geometry(mapped, stencil.faces, heat_fwhm)
    depth = depth_potential(smoothed, stencil.faces, solver=solver, max_iter=max_iter)
    vals = O.smooth_values(depth.values, smoothed, stencil.faces, 50.0)
    lo, hi = vals.min(), vals.max()
    return (vals - lo) / max(hi - lo, _EPS64)


# ---------------------------------------------------------------- rotation search


def grid_index(points, faces, grid_size: int = 128, margin: int = 1):
    """CSR of triangles per cell of a grid on [-1, 1]^3 (surface_rotation.py:57-120):
    the cells of each face's bounding box (± margin, clamped), x fastest, then y,
    then z; faces ascending within a cell.  Vectorised form of oracle.grid_index."""

    points = np.asarray(points, dtype=np.float64)
    faces = np.asarray(faces, dtype=np.int64)
    cell = 2.0 / float(grid_size)
    tri = points[faces]
    lo = np.floor((tri.min(axis=1) + 1.0) / cell).astype(np.int64)
    hi = np.floor((tri.max(axis=1) + 1.0) / cell).astype(np.int64)
    lo = np.clip(lo - margin, 0, grid_size - 1)
    hi = np.clip(hi + margin, 0, grid_size - 1)
    ext = hi - lo + 1
    per_face = ext[:, 0] * ext[:, 1] * ext[:, 2]
    owners = np.repeat(np.arange(faces.shape[0], dtype=np.int64), per_face)
    start = np.concatenate([[0], np.cumsum(per_face)[:-1]])
    k = np.arange(owners.size, dtype=np.int64) - start[owners]
    ex, ey = ext[owners, 0], ext[owners, 1]
    x = lo[owners, 0] + k % ex
    y = lo[owners, 1] + (k // ex) % ey
    z = lo[owners, 2] + k // (ex * ey)
    cells = x + grid_size * (y + grid_size * z)
    order = np.lexsort((owners, cells))
    counts = np.bincount(cells, minlength=grid_size ** 3)
    off = np.zeros(grid_size ** 3 + 1, np.int32)
    off[1:] = np.cumsum(counts)
    return off, owners[order].astype(np.int32)


@dataclass
class RotationResult:
    angle: np.ndarray
    cost: float
    seed_index: int
    seed_costs: np.ndarray
    iterations: int


def rotation_cost(target: Stencil, source_points, source_values, target_values, angles, *,
                  grid_size: int = 128, margin: int = 1) -> np.ndarray:
    """The reference's rotation SSD (surface_rotation.py:512-546) of angle triples."""

    index = grid_index(target.sphere64, target.faces, grid_size, margin)
    return O.rotation_costs(np.asarray(source_points, dtype=np.float32), source_values,
                            target_values, target.sphere64, target.faces,
                            np.asarray(angles, dtype=np.float64).reshape(-1, 3), grid_size,
                            margin, index=index)


def search_rotation(target: Stencil, source_points, source_values, target_values, *,
                    grid_size: int = 128, margin: int = 1, refine: bool = True,
                    max_iter: int = 500, tol: float = 1e-4,
                    simplex_step: float = 0.1) -> RotationResult:
    """46-seed scan, first minimum, then CAT-ordered Nelder-Mead
    (surface_rotation.py:562-699, gpu_pipeline.py:908-919)."""

    index = grid_index(target.sphere64, target.faces, grid_size, margin)
    src = np.asarray(source_points, dtype=np.float32)

    def cost(angles):
        return O.rotation_costs(src, source_values, target_values, target.sphere64,
                                target.faces, angles, grid_size, margin, index=index)

    seeds = O.official_seed_grid_angles()
    seed_costs = cost(seeds)
    best = int(np.argmin(seed_costs))
    if not refine:
        return RotationResult(seeds[best].copy(), float(seed_costs[best]), best, seed_costs, 0)
    angle, c, iterations = O.nelder_mead(cost, seeds[best], max_iter=max_iter, tol=tol,
                                         simplex_step=simplex_step)
    return RotationResult(angle, float(c), best, seed_costs, int(iterations))


# ---------------------------------------------------------------- DARTEL side


def curvature_maps(surface_vertices, stencil: Stencil, fwhms, curvtypes=None) -> np.ndarray:
    """[steps, ny, nx] curvature sheets of one surface: build_maps of
    surface_warp.py:427-448 (several all-type-5 steps together,
    surface_stencil.py:521-550; else one curvature_to_sheet per step, :624-644)."""

    fwhms = tuple(fwhms)
    types = tuple(int(t) for t in (curvtypes if curvtypes is not None else (5,) * len(fwhms)))
    geometry = O.resample_vertices_f32(surface_vertices, stencil.surf_idx, stencil.surf_w)
    if len(types) > 1 and all(t == 5 for t in types):
        return np.stack(O.curvature_type5_to_sheet_many(geometry, stencil.faces,
                                                        stencil.sheet_idx, stencil.sheet_w,
                                                        fwhms))
    return np.stack([O.curvature_to_sheet(geometry, stencil.faces, stencil.sheet_idx,
                                          stencil.sheet_w, f, t) for f, t in zip(fwhms, types)])


def solve_from_surfaces(source_vertices, target_vertices, source: Stencil, target: Stencil,
                        fwhms, *, curvtypes=None, loop: int = 6, cycles: int = 3, nit: int = 3,
                        its: int = 3, code: int = 1):
    """Sheets of both sides and their DARTEL flow (surface_warp.py:370-488)."""

    src = curvature_maps(source_vertices, source, fwhms, curvtypes)
    tgt = curvature_maps(target_vertices, target, fwhms, curvtypes)
    flow, metrics = O.solve_dartel_maps(src, tgt, loop=loop, lmreg=1e-3, cycles=cycles,
                                        nit=nit, its=its, code=code)
    return flow, metrics, src, tgt


def warp(stencil: Stencil, flow) -> np.ndarray:
    """Inverse warp of the stencil's unit points (surface_warp.py:587-617), float64."""

    return O.apply_flow_to_sphere(stencil.unit_points, flow, inverse=True)


# ---------------------------------------------------------------- whole registration


@dataclass
class SurfWarpResult:
    vertices: np.ndarray              # [N, 3] float32, the output sphere
    rotation: RotationResult
    features: tuple[np.ndarray, np.ndarray]
    flows: list = field(default_factory=list)
    stage_seconds: dict = field(default_factory=dict)
    stencils: tuple = ()              # (source, target) stencils of the input spheres


def surfwarp(source_surface, source_sphere, source_faces, target_surface, target_sphere,
             target_faces, *, steps: int = 2, runs: int = 2, avg: bool = True,
             fwhm=(5.0, 5.0 / 3.0, 5.0 / 9.0), curvtypes=(5, 5, 2), loop: int = 6, cycles: int = 3, nit: int = 3,
             its: int = 3, code: int = 1, map_size=(512, 256), features=None,
             rotation_grid: int = 128, rotation_margin: int = 1, max_iter: int = 500,
             tol: float = 1e-4, simplex_step: float = 0.1,
             feature_solver: str = "colored-sor") -> SurfWarpResult:
    """One CAT_SurfWarp registration (gpu_pipeline.py:468-1326) on the CPU oracle.

    ``features`` = (source values, target values) on the resampled spheres skips the
    rotation-feature stage (the reference's rotation_values_path)."""

    nx, ny = (int(s) for s in map_size)
    types = tuple(int(t) for t in curvtypes[:steps])
    stage: dict[str, float] = {}

    def timed(name, fn, *a, **k):
        t = time.perf_counter()
        out = fn(*a, **k)
        stage[name] = stage.get(name, 0.0) + time.perf_counter() - t
        return out

    src_faces, tgt_faces = O._ci(source_faces), O._ci(target_faces)
    src_surf = O._c32(source_surface)
    tgt_surf = O._c32(target_surface)
    src_st = timed("stencils", build_stencil, source_sphere, src_faces, normalise=True, nx=nx,
                   ny=ny)
    tgt_st = timed("stencils", build_stencil, target_sphere, tgt_faces, normalise=True, nx=nx,
                   ny=ny)
    if features is None:
        src_vals = timed("rotation_features", rotation_feature, src_st, src_surf, 15.0,
                         solver=feature_solver)
        tgt_vals = timed("rotation_features", rotation_feature, tgt_st, tgt_surf, 10.0,
                         solver=feature_solver)
    else:
        src_vals, tgt_vals = (np.asarray(f, dtype=np.float64) for f in features)
    rot = timed("rotation_search", search_rotation, tgt_st, src_st.sphere64, src_vals, tgt_vals,
                grid_size=rotation_grid, margin=rotation_margin, max_iter=max_iter, tol=tol,
                simplex_step=simplex_step)
    current = O.rotate_vertices(O.normalise_sphere(source_sphere), rot.angle)
    avg_tgt_surface = O.rotate_vertices(tgt_surf, POLE)
    avg_tgt_sphere = O.rotate_vertices(O.normalise_sphere(target_sphere), POLE)
    cur_st = timed("stencils", build_stencil, current, src_faces, normalise=False, nx=nx, ny=ny)
    avg_tgt_st = (timed("stencils", build_stencil, avg_tgt_sphere, tgt_faces, normalise=False,
                        nx=nx, ny=ny) if avg else None)
    flows = []
    last_flow = prewarp = avg_src_surface = avg_src_st = None
    for run in range(runs):
        run_fwhm = tuple(float(fwhm[0]) / (3.0 ** (steps * (run + 1) + i)) for i in range(steps))
        final = avg and run == runs - 1
        if final:
            prewarp = current.copy()
            avg_src_surface = O.rotate_vertices(src_surf, POLE)
            avg_src_sphere = O.rotate_vertices(prewarp, POLE)
            avg_src_st = timed("stencils", build_stencil, avg_src_sphere, src_faces,
                               normalise=False, nx=nx, ny=ny)
        flow, _, _, _ = timed("dartel", solve_from_surfaces, src_surf, tgt_surf, cur_st, tgt_st,
                              run_fwhm, curvtypes=types, loop=loop, cycles=cycles, nit=nit,
                              its=its, code=code)
        flows.append(flow)
        if not final:
            current = timed("warp", warp, cur_st, flow).astype(np.float32)
        last_flow = flow
        if run + 1 < runs:
            cur_st = timed("stencils", build_stencil, current, src_faces, normalise=False,
                           nx=nx, ny=ny)
    if avg:
        avg_fwhm = tuple(float(fwhm[0]) / (3.0 ** (steps * (runs + 1) + i)) for i in range(steps))
        avg_flow, _, _, _ = timed("dartel", solve_from_surfaces, avg_src_surface,
                                  avg_tgt_surface, avg_src_st, avg_tgt_st, avg_fwhm,
                                  curvtypes=types, loop=loop, cycles=cycles, nit=nit, its=its,
                                  code=code)
        flows.append(avg_flow)
        main_warped = timed("warp", warp, cur_st, last_flow)
        avg_warped = timed("warp", warp, avg_src_st, avg_flow)
        unrotate = O.rotation_matrix_from_angles(-POLE)
        current = timed("warp", O.average_xz_surfaces, avg_warped @ unrotate.T, main_warped)
    return SurfWarpResult(vertices=np.ascontiguousarray(current, dtype=np.float32),
                          rotation=rot, features=(src_vals, tgt_vals), flows=flows,
                          stage_seconds=stage, stencils=(src_st, tgt_st))


def timed_surfwarp(*args, **kwargs):
    """surfwarp with its wall time and this process's user + system CPU time
    (getrusage, as tools/audit_cat_warpsurf.py:52-60 records the CPU reference)."""

    r0 = resource.getrusage(resource.RUSAGE_SELF)
    t0 = time.perf_counter()
    result = surfwarp(*args, **kwargs)
    wall = time.perf_counter() - t0
    r1 = resource.getrusage(resource.RUSAGE_SELF)
    cpu = (r1.ru_utime - r0.ru_utime) + (r1.ru_stime - r0.ru_stime)
    return result, wall, cpu
