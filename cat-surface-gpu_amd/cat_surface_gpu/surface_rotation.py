"""CAT initial rotation: grid index, SSD cost and search, on MI355X.

Drop-in for src/cat_surface_gpu/surface_rotation.py.  The candidate index is
built on the GPU; every cost evaluation (rotate in float32, locate in the
target triangle index, barycentric sample of the target feature, float64
SSD) is one HIP launch over all points × all candidate angles.  The 46-seed
grid and the CAT-ordered Nelder-Mead follow the reference step for step; the
simplex bookkeeping runs on the host with one cost launch per evaluation.
"""

from __future__ import annotations

import ctypes
import functools
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, schedule
from .locate import _CAPACITY, GridIndex, build_grid_index, build_grid_index_async


def official_seed_grid_angles() -> np.ndarray:
    """The 46 upstream seeds, identity first (surface_rotation.py:14-23)."""

    seed0 = (-0.6, -0.3, 0.0, 0.3, 0.6)
    seed12 = (-0.3, 0.0, 0.3)
    values = [(0.0, 0.0, 0.0)]
    values.extend((a, b, c) for a in seed0 for b in seed12 for c in seed12)
    return np.asarray(values, dtype=np.float64)


def rotation_matrix_np(angles) -> np.ndarray:
    """Rz (Ry Rx) assembled column-major then read row-major (float64)."""

    a, b, g = (float(x) for x in np.asarray(angles, dtype=np.float64).reshape(3))
    ca, sa, cb, sb, cg, sg = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(g), np.sin(g)
    rx = [1.0, 0.0, 0.0, 0.0, ca, sa, 0.0, -sa, ca]
    ry = [cb, 0.0, sb, 0.0, 1.0, 0.0, -sb, 0.0, cb]
    rz = [cg, sg, 0.0, -sg, cg, 0.0, 0.0, 0.0, 1.0]
    first = [0.0] * 9
    res = [0.0] * 9
    for r in range(3):
        for c in range(3):
            first[r + 3 * c] = sum(ry[r + 3 * i] * rx[i + 3 * c] for i in range(3))
    for r in range(3):
        for c in range(3):
            res[r + 3 * c] = sum(rz[r + 3 * i] * first[i + 3 * c] for i in range(3))
    return np.asarray(res, dtype=np.float64).reshape(3, 3)


def angles_from_matrix(matrix) -> np.ndarray:
    """Euler angles (alpha, beta, gamma) with rotation_matrix_np(angles) == matrix.

    rotation_matrix_np returns M = (Rz(gamma) Ry(-beta) Rx(alpha))ᵀ, so
    M[0, 2] = sin(beta), M[1, 2] / M[2, 2] = tan(alpha) and
    M[0, 1] / M[0, 0] = tan(gamma) (cos(beta) > 0 branch).
    """

    m = np.asarray(matrix, dtype=np.float64).reshape(3, 3)
    beta = float(np.arcsin(np.clip(m[0, 2], -1.0, 1.0)))
    alpha = float(np.arctan2(m[1, 2], m[2, 2]))
    gamma = float(np.arctan2(m[0, 1], m[0, 0]))
    return np.array([alpha, beta, gamma], dtype=np.float64)


def quaternion_matrix(q) -> np.ndarray:
    """Rotation matrices [..., 3, 3] of unit quaternions [..., 4] (w, x, y, z)."""

    q = np.asarray(q, dtype=np.float64)
    q = q / np.linalg.norm(q, axis=-1, keepdims=True)
    w, x, y, z = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1),
    ], -2)


def hopf_grid(n_sphere: int = 256, n_circle: int = 32) -> np.ndarray:
    """n_sphere x n_circle rotation matrices spread uniformly over SO(3).

    Hopf-fibration sampling (Yershova, Jain, LaValle and Mitchell, "Generating
    uniform incremental grids on SO(3) using the Hopf fibration", IJRR 2010):
    a near-uniform set of base points (theta, phi) on S² (a Fibonacci lattice
    here) times n_circle equally spaced fibre angles psi in [0, 2 pi), mapped to
    the unit quaternion (cos(theta/2) cos(psi/2), cos(theta/2) sin(psi/2),
    sin(theta/2) cos(phi + psi/2), sin(theta/2) sin(phi + psi/2)).  The map
    carries the product measure on S² x S¹ to the uniform measure on SO(3).
    Candidate c is base point c // n_circle, fibre angle c % n_circle.
    """

    if n_sphere < 1 or n_circle < 1:
        raise ValueError("hopf_grid needs at least one base point and one fibre angle")
    k = np.arange(n_sphere, dtype=np.float64)
    z = 1.0 - (2.0 * k + 1.0) / n_sphere
    theta = np.arccos(np.clip(z, -1.0, 1.0))
    phi = np.mod(k * np.pi * (3.0 - np.sqrt(5.0)), 2.0 * np.pi)
    psi = (np.arange(n_circle, dtype=np.float64) + 0.5) * (2.0 * np.pi / n_circle)
    th = np.repeat(theta, n_circle)
    ph = np.repeat(phi, n_circle)
    ps = np.tile(psi, n_sphere)
    q = np.stack([np.cos(th / 2) * np.cos(ps / 2), np.cos(th / 2) * np.sin(ps / 2),
                  np.sin(th / 2) * np.cos(ph + ps / 2), np.sin(th / 2) * np.sin(ph + ps / 2)],
                 axis=-1)
    return quaternion_matrix(q)


def axis_angle_matrix(delta) -> np.ndarray:
    """exp([delta]x) by Rodrigues' formula."""

    d = np.asarray(delta, dtype=np.float64).reshape(3)
    theta = float(np.linalg.norm(d))
    k = np.array([[0.0, -d[2], d[1]], [d[2], 0.0, -d[0]], [-d[1], d[0], 0.0]])
    if theta < 1e-12:
        return np.eye(3) + k
    return (np.eye(3) + (np.sin(theta) / theta) * k
            + ((1.0 - np.cos(theta)) / (theta * theta)) * (k @ k))


@dataclass(frozen=True)
class RotationGridIndex:
    """Host view of the candidate index (surface_rotation.py:26-180)."""

    points: np.ndarray
    faces: np.ndarray
    candidate_table: np.ndarray | None
    grid_size: int
    margin: int
    candidate_offsets: np.ndarray | None = None
    candidate_faces: np.ndarray | None = None

    @classmethod
    def from_stencil(cls, stencil, *, grid_size: int = 128, margin: int = 1,
                     include_dense_candidate_table: bool = False) -> "RotationGridIndex":
        return cls.from_geometry(stencil.sphere_points, stencil.faces, grid_size=grid_size,
                                 margin=margin,
                                 include_dense_candidate_table=include_dense_candidate_table)

    @classmethod
    def from_geometry(cls, points, faces, *, grid_size: int = 128, margin: int = 1,
                      include_dense_candidate_table: bool = False,
                      device: str | torch.device = "auto") -> "RotationGridIndex":
        """Build on the GPU and keep host copies of the CSR arrays."""

        from .dartel_grid import resolve_device

        points = np.asarray(points, dtype=np.float64)
        faces = np.asarray(faces, dtype=np.int32)
        if points.ndim != 2 or points.shape[1] != 3:
            raise ValueError("points must have shape [n, 3]")
        if faces.ndim != 2 or faces.shape[1] != 3:
            raise ValueError("faces must have shape [m, 3]")
        if grid_size < 4 or margin < 0:
            raise ValueError("grid_size must be at least 4 and margin must be non-negative")
        dev = resolve_device(device)
        index = build_grid_index(torch.as_tensor(points, device=dev),
                                 torch.as_tensor(faces, device=dev), grid_size, margin)
        offsets = index.cell_off.cpu().numpy()
        cell_faces = index.cell_faces.cpu().numpy()
        table = None
        if include_dense_candidate_table:
            width = int(np.diff(offsets).max())
            table = np.full((grid_size ** 3, width), -1, dtype=np.int32)
            for cell in np.flatnonzero(np.diff(offsets)):
                vals = cell_faces[offsets[cell]:offsets[cell + 1]]
                table[cell, : vals.size] = vals
        return cls(points=np.ascontiguousarray(points), faces=np.ascontiguousarray(faces),
                   candidate_table=table, grid_size=int(grid_size), margin=int(margin),
                   candidate_offsets=offsets, candidate_faces=cell_faces)

    @property
    def max_candidates(self) -> int:
        if self.candidate_offsets is not None:
            return int(np.diff(self.candidate_offsets).max())
        return 0 if self.candidate_table is None else int(self.candidate_table.shape[1])

    def to(self, device, *, geometry_dtype: torch.dtype = torch.float32,
           candidate_table_dtype: torch.dtype = torch.int32,
           compressed_candidate_table: bool = True) -> "RotationGridIndexDevice":
        from .dartel_grid import resolve_device

        dev = resolve_device(device)
        if candidate_table_dtype not in (torch.int32, torch.int64):
            raise ValueError("candidate_table_dtype must be int32 or int64")
        triangles = self.points[self.faces]
        return RotationGridIndexDevice(
            target_points=torch.as_tensor(self.points, dtype=geometry_dtype, device=dev),
            target_faces=torch.as_tensor(self.faces, dtype=torch.int32, device=dev),
            target_triangles=torch.as_tensor(triangles, dtype=torch.float32, device=dev),
            grid_size=self.grid_size,
            candidate_offsets=torch.as_tensor(self.candidate_offsets, device=dev),
            candidate_faces=torch.as_tensor(self.candidate_faces, device=dev),
        )


@dataclass(frozen=True)
class RotationCostInputs:
    source_points: torch.Tensor
    source_values: torch.Tensor
    target_values: torch.Tensor


@dataclass(frozen=True)
class RotationSearchResult:
    seed_angle: torch.Tensor
    seed_cost: torch.Tensor
    seed_costs: torch.Tensor
    angle: torch.Tensor
    cost: torch.Tensor
    rotation_matrix: torch.Tensor
    iterations: int


# The rotation locator skips candidates by their faces' bounding spheres
# (identical results); False scans every candidate of a cell.
_LOCATOR_SKIP = True


def set_locator_skip(enabled: bool) -> bool:
    """Select the bounding-sphere candidate skip of the rotation locator; returns
    the previous setting."""

    global _LOCATOR_SKIP
    previous = _LOCATOR_SKIP
    _LOCATOR_SKIP = bool(enabled)
    return previous


@dataclass
class RotationGridIndexDevice:
    """Evaluate CAT rotation candidates against a device-resident index."""

    target_points: torch.Tensor
    target_faces: torch.Tensor
    target_triangles: torch.Tensor
    grid_size: int
    candidate_offsets: torch.Tensor
    candidate_faces: torch.Tensor
    candidate_table: torch.Tensor | None = None

    @classmethod
    def from_device_geometry(cls, points64: torch.Tensor, faces: torch.Tensor, *,
                             grid_size: int = 128, margin: int = 1) -> "RotationGridIndexDevice":
        """Build straight from device tensors (no host round trip)."""

        f = faces.to(torch.int32).contiguous()
        if schedule.capturing():
            # the pair capacity learned by earlier builds; the count the index
            # needed is checked after the replay
            key = (int(f.shape[0]), int(grid_size), int(margin))
            if key not in _CAPACITY:
                raise RuntimeError("no learned index capacity for this mesh size; run the "
                                   "schedule eagerly first")
            required = torch.zeros(1, dtype=torch.int32, device=f.device)
            index = build_grid_index_async(points64.to(torch.float64), f, grid_size, margin,
                                           _CAPACITY[key], required)
            schedule.defer(schedule._exceeds([required], _CAPACITY[key]),
                           "rotation index capacity")
        else:
            index = build_grid_index(points64.to(torch.float64), f, grid_size, margin)
        return cls(target_points=points64.to(torch.float32).contiguous(), target_faces=f,
                   target_triangles=points64[f.long()].to(torch.float32).contiguous(),
                   grid_size=grid_size, candidate_offsets=index.cell_off,
                   candidate_faces=index.cell_faces)

    @property
    def device(self) -> torch.device:
        return self.target_points.device

    def face_spheres(self) -> torch.Tensor | None:
        """The target faces' bounding spheres ([F, 4] float32, built once) that let
        the locator skip candidates (csg_face_spheres); None when skipping is off."""

        if not _LOCATOR_SKIP:
            return None
        sph = self.__dict__.get("_spheres")
        if sph is None:
            tri = self.target_triangles.contiguous()
            nf = int(tri.shape[0])
            sph = torch.empty((nf, 4), dtype=torch.float32, device=self.device)
            _lib.check(_lib.lib().csg_face_spheres(_lib.ptr(tri), nf, _lib.ptr(sph),
                                                   _lib.stream_handle(self.device)),
                       "csg_face_spheres")
            self.__dict__["_spheres"] = sph
        return sph

    def rotation_matrix(self, angles) -> torch.Tensor:
        a = torch.as_tensor(angles, dtype=torch.float64).reshape(-1)
        if a.numel() != 3:
            raise ValueError("rotation_matrix accepts exactly one angle vector")
        return torch.as_tensor(rotation_matrix_np(a.cpu().numpy()), dtype=torch.float32,
                               device=self.device)

    def rotate_points(self, source_points, angles) -> torch.Tensor:
        src = torch.as_tensor(source_points, dtype=torch.float32, device=self.device)
        ang = torch.as_tensor(angles, dtype=torch.float64).reshape(-1, 3).cpu().numpy()
        mats = torch.as_tensor(np.stack([rotation_matrix_np(a) for a in ang]),
                               dtype=torch.float32, device=self.device)
        return torch.matmul(src.unsqueeze(0), mats.transpose(1, 2))

    def prepare_cost_inputs(self, source_points, source_values, target_values) -> RotationCostInputs:
        src = torch.as_tensor(source_points, dtype=torch.float32, device=self.device).contiguous()
        sv = torch.as_tensor(source_values, dtype=torch.float64, device=self.device).reshape(-1)
        tv = torch.as_tensor(target_values, dtype=torch.float64, device=self.device).reshape(-1)
        if src.ndim != 2 or src.shape[1] != 3:
            raise ValueError("source_points must have shape [n, 3]")
        if sv.numel() != src.shape[0]:
            raise ValueError("source_values length does not match source_points")
        if tv.numel() != self.target_points.shape[0]:
            raise ValueError("target_values length does not match the target sphere")
        return RotationCostInputs(src, sv.contiguous(), tv.contiguous())

    def _launch_costs(self, inputs: RotationCostInputs, angles, locate: bool = False,
                      uncovered_dev: torch.Tensor | None = None):
        """Enqueue one cost launch; with ``uncovered_dev`` nothing synchronises."""

        ang = torch.as_tensor(angles, dtype=torch.float64, device=self.device).reshape(-1, 3)
        ang = ang.contiguous()
        na, ns = int(ang.shape[0]), int(inputs.source_points.shape[0])
        costs = torch.empty(na, dtype=torch.float64, device=self.device)
        loc_f = loc_b = None
        if locate:
            loc_f = torch.empty((na, ns), dtype=torch.int32, device=self.device)
            loc_b = torch.empty((na, ns, 3), dtype=torch.float32, device=self.device)
        lib = _lib.lib()
        ws = _lib.workspace(lib.csg_rotation_costs_workspace_bytes(ns, na), self.device)
        uncovered = ctypes.c_int()
        _lib.check(lib.csg_rotation_costs(
            _lib.ptr(inputs.source_points), _lib.ptr(inputs.source_values), ns,
            _lib.ptr(inputs.target_values), _lib.ptr(self.target_faces),
            _lib.ptr(self.target_triangles), _lib.ptr(self.face_spheres()),
            _lib.ptr(self.candidate_offsets), _lib.ptr(self.candidate_faces),
            int(self.grid_size), _lib.ptr(ang), na,
            _lib.ptr(costs), _lib.ptr(loc_f), _lib.ptr(loc_b), _lib.ptr(ws), ws.numel(),
            None if uncovered_dev is not None else ctypes.byref(uncovered),
            _lib.ptr(uncovered_dev), _lib.stream_handle(self.device)), "csg_rotation_costs")
        return costs, loc_f, loc_b, (None if uncovered_dev is not None else uncovered.value)

    def _launch_matrix_costs(self, inputs: RotationCostInputs, matrices, gradient: bool,
                             uncovered_dev: torch.Tensor) -> torch.Tensor:
        """Enqueue the matrix-input cost ([na]) or cost + gradient + Gauss-Newton
        terms ([na, 10]); the uncovered count goes to ``uncovered_dev``."""

        if isinstance(matrices, torch.Tensor):
            mats = matrices.to(device=self.device, dtype=torch.float32).reshape(-1, 9)
        else:
            mats = torch.as_tensor(np.asarray(matrices, dtype=np.float64).reshape(-1, 9),
                                   dtype=torch.float32).to(self.device)
        mats = mats.contiguous()
        na, ns = int(mats.shape[0]), int(inputs.source_points.shape[0])
        out = torch.empty((na, 10) if gradient else (na,), dtype=torch.float64,
                          device=self.device)
        lib = _lib.lib()
        ws = _lib.workspace(lib.csg_rotation_cost_matrices_workspace_bytes(ns, na,
                                                                           int(gradient)),
                            self.device)
        _lib.check(lib.csg_rotation_cost_matrices(
            _lib.ptr(inputs.source_points), _lib.ptr(inputs.source_values), ns,
            _lib.ptr(inputs.target_values), _lib.ptr(self.target_faces),
            _lib.ptr(self.target_triangles), _lib.ptr(self.face_spheres()),
            _lib.ptr(self.candidate_offsets), _lib.ptr(self.candidate_faces),
            int(self.grid_size), _lib.ptr(mats), na,
            int(gradient), _lib.ptr(out), _lib.ptr(ws), ws.numel(), _lib.ptr(uncovered_dev),
            _lib.stream_handle(self.device)), "csg_rotation_cost_matrices")
        return out

    def cost_matrices(self, inputs: RotationCostInputs, matrices, *, gradient: bool = False):
        """SSD per candidate rotation matrix; with ``gradient`` also the axis-angle
        gradient [na, 3] and Gauss-Newton matrix [na, 3, 3] (float64 tensors)."""

        unc = torch.zeros(1, dtype=torch.int32, device=self.device)
        out = self._launch_matrix_costs(inputs, matrices, gradient, unc)
        if int(unc.item()):
            raise RuntimeError("Rotation is not covered by target sphere candidate index")
        if not gradient:
            return out
        return out[:, 0], out[:, 1:4], _gn_matrix(out[:, 4:10])

    def _costs(self, inputs: RotationCostInputs, angles: torch.Tensor,
               locate: bool = False):
        costs, loc_f, loc_b, uncovered = self._launch_costs(inputs, angles, locate)
        if uncovered:
            raise RuntimeError("Rotation is not covered by target sphere candidate index")
        return costs, loc_f, loc_b

    def cost_batch_prepared(self, inputs: RotationCostInputs, angles, *,
                            point_chunk: int = 1024) -> torch.Tensor:
        """SSD per candidate angle (surface_rotation.py:512-546).

        ``point_chunk`` is accepted for API compatibility; all points of all
        candidates go to the device in one launch.
        """

        if point_chunk < 1:
            raise ValueError("point_chunk must be positive")
        return self._costs(inputs, angles)[0]

    def cost_batch(self, source_points, source_values, target_values, angles, *,
                   point_chunk: int = 1024) -> torch.Tensor:
        inputs = self.prepare_cost_inputs(source_points, source_values, target_values)
        return self.cost_batch_prepared(inputs, angles, point_chunk=point_chunk)

    def locate(self, inputs: RotationCostInputs, angles):
        """(faces [na, ns], barycentric weights [na, ns, 3]) of the rotated points."""

        _, f, b = self._costs(inputs, angles, locate=True)
        return f, b

    def coarse_seed_search_prepared(self, inputs: RotationCostInputs, *, point_chunk: int = 1024):
        """All 46 seeds, first minimum (surface_rotation.py:562-575)."""

        angles = torch.as_tensor(official_seed_grid_angles(), dtype=torch.float64,
                                 device=self.device)
        costs = self.cost_batch_prepared(inputs, angles, point_chunk=point_chunk)
        best = int(torch.argmin(costs).item())
        return angles[best], costs[best], costs

    def coarse_seed_search(self, source_points, source_values, target_values, *,
                           point_chunk: int = 1024):
        inputs = self.prepare_cost_inputs(source_points, source_values, target_values)
        return self.coarse_seed_search_prepared(inputs, point_chunk=point_chunk)

    def refine_nelder_mead_prepared(self, inputs: RotationCostInputs, initial_angles, *,
                                    point_chunk: int = 1024, max_iter: int = 500,
                                    tol: float = 1.0e-4, simplex_step: float = 0.1):
        """CAT-ordered Nelder-Mead (surface_rotation.py:594-699).

        The simplex bookkeeping is host float64 exactly as the reference's;
        each evaluation is one device cost launch.
        """

        return refine_nelder_mead_many([(self, inputs, initial_angles)], max_iter=max_iter,
                                       tol=tol, simplex_step=simplex_step)[0]

    def refine_nelder_mead(self, source_points, source_values, target_values, initial_angles, *,
                           point_chunk: int = 1024, max_iter: int = 500, tol: float = 1.0e-4,
                           simplex_step: float = 0.1):
        inputs = self.prepare_cost_inputs(source_points, source_values, target_values)
        return self.refine_nelder_mead_prepared(inputs, initial_angles, point_chunk=point_chunk,
                                                max_iter=max_iter, tol=tol,
                                                simplex_step=simplex_step)

    def search_rotation_prepared(self, inputs: RotationCostInputs, *, point_chunk: int = 1024,
                                 refine: bool = True, max_iter: int = 500, tol: float = 1.0e-4,
                                 simplex_step: float = 0.1) -> RotationSearchResult:
        seed_angle, seed_cost, seed_costs = self.coarse_seed_search_prepared(
            inputs, point_chunk=point_chunk)
        if refine:
            angle, cost, iterations = self.refine_nelder_mead_prepared(
                inputs, seed_angle, point_chunk=point_chunk, max_iter=max_iter, tol=tol,
                simplex_step=simplex_step)
        else:
            angle, cost, iterations = seed_angle.clone(), seed_cost.clone(), 0
        return RotationSearchResult(seed_angle=seed_angle.clone(), seed_cost=seed_cost.clone(),
                                    seed_costs=seed_costs, angle=angle, cost=cost,
                                    rotation_matrix=self.rotation_matrix(angle),
                                    iterations=iterations)

    def search_rotation(self, source_points, source_values, target_values, *,
                        point_chunk: int = 1024, refine: bool = True, max_iter: int = 500,
                        tol: float = 1.0e-4, simplex_step: float = 0.1) -> RotationSearchResult:
        inputs = self.prepare_cost_inputs(source_points, source_values, target_values)
        return self.search_rotation_prepared(inputs, point_chunk=point_chunk, refine=refine,
                                             max_iter=max_iter, tol=tol,
                                             simplex_step=simplex_step)


def _nelder_mead_steps(seed, max_iter: int, tol: float, simplex_step: float):
    """CAT-ordered Nelder-Mead (surface_rotation.py:594-699) as a generator: it
    yields the angle triples it needs evaluated (an [k, 3] array) and receives
    their costs; it returns (angle, cost, iterations)."""

    simplex = np.stack([seed, seed + (simplex_step, 0.0, 0.0), seed + (0.0, simplex_step, 0.0),
                        seed + (0.0, 0.0, simplex_step)])
    f = np.array((yield simplex.copy()), dtype=np.float64)
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
        fr = (yield refl[None])[0]
        if fr < f[lo]:
            exp = centroid + 2.0 * (refl - centroid)
            fe = (yield exp[None])[0]
            if fe < fr:
                simplex[hi], f[hi] = exp, fe
            else:
                simplex[hi], f[hi] = refl, fr
        elif fr < f[sh]:
            simplex[hi], f[hi] = refl, fr
        else:
            con = centroid + 0.5 * (simplex[hi] - centroid)
            fc = (yield con[None])[0]
            if fc < f[hi]:
                simplex[hi], f[hi] = con, fc
            else:
                for i in range(4):
                    if i != lo:
                        simplex[i] = simplex[lo] + 0.5 * (simplex[i] - simplex[lo])
                        f[i] = (yield simplex[i][None])[0]
        if np.max(np.abs(f - f[lo])) < tol:
            break
    return simplex[lo].copy(), float(f[lo]), iteration + 1


def refine_nelder_mead_many(problems, *, max_iter: int = 500, tol: float = 1.0e-4,
                            simplex_step: float = 0.1):
    """Nelder-Mead refinement of several independent rotation problems in
    lockstep.  ``problems`` holds (index, prepared inputs, seed angles); each
    round launches the cost evaluations every problem's simplex needs and reads
    all costs back with one synchronisation.  Each problem follows exactly the
    sequential algorithm; returns [(angle, cost, iterations)] as tensors."""

    if max_iter < 1 or tol < 0.0 or simplex_step <= 0.0:
        raise ValueError("max_iter, tol, and simplex_step must be positive")
    gens, pending, results = [], [], [None] * len(problems)
    for k, (index, inputs, initial) in enumerate(problems):
        seed = np.asarray(torch.as_tensor(initial, dtype=torch.float64).cpu(),
                          dtype=np.float64).reshape(-1)
        if seed.size != 3:
            raise ValueError("initial_angles must contain exactly three angles")
        gen = _nelder_mead_steps(seed, max_iter, tol, simplex_step)
        gens.append(gen)
        pending.append(next(gen))
    active = list(range(len(problems)))
    while active:
        launched = []
        for k in active:
            index, inputs, _ = problems[k]
            unc = torch.zeros(1, dtype=torch.int32, device=index.device)
            costs, _, _, _ = index._launch_costs(inputs, torch.as_tensor(pending[k]),
                                                 uncovered_dev=unc)
            launched.append((k, costs, unc))
        host = torch.cat([torch.cat([c.to(torch.float64),
                                     u.to(torch.float64).to(c.device)]) for _, c, u in launched])
        host = host.cpu().numpy()
        still, pos = [], 0
        for k, costs, _ in launched:
            n = int(costs.numel())
            vals, unc = host[pos:pos + n], host[pos + n]
            pos += n + 1
            if unc:
                raise RuntimeError("Rotation is not covered by target sphere candidate index")
            try:
                pending[k] = gens[k].send(vals)
                still.append(k)
            except StopIteration as done:
                angle, cost, iterations = done.value
                dev = problems[k][0].device
                results[k] = (torch.as_tensor(angle, device=dev),
                              torch.as_tensor(cost, dtype=torch.float64, device=dev), iterations)
        active = still
    return results


def _gn_matrix(terms):
    """[..., 6] (xx, yy, zz, xy, xz, yz) -> symmetric [..., 3, 3]."""

    t = terms
    if isinstance(t, torch.Tensor):
        rows = [torch.stack([t[..., 0], t[..., 3], t[..., 4]], -1),
                torch.stack([t[..., 3], t[..., 1], t[..., 5]], -1),
                torch.stack([t[..., 4], t[..., 5], t[..., 2]], -1)]
        return torch.stack(rows, -2)
    t = np.asarray(t)
    return np.stack([np.stack([t[..., 0], t[..., 3], t[..., 4]], -1),
                     np.stack([t[..., 3], t[..., 1], t[..., 5]], -1),
                     np.stack([t[..., 4], t[..., 5], t[..., 2]], -1)], -2)


def refine_adjoint_launch(index, inputs, start: torch.Tensor, *, start_is_angles: bool,
                          steps: int = 20, max_step: float = 0.05, damping: float = 1e-3):
    """Enqueue csg_rotation_refine for one problem (nothing synchronises).

    ``start`` is a device tensor of m Euler-angle triples ([m, 3],
    ``start_is_angles``) or m row-major matrices ([m, 9] / [m, 3, 3]).
    Returns device tensors (matrices [m, 3, 3], angles [m, 3], costs [m] in the
    reference's angle formulation, uncovered-point count [1]).
    """

    st = start.to(device=index.device, dtype=torch.float64).contiguous()
    m = int(st.shape[0])
    ns = int(inputs.source_points.shape[0])
    mats = torch.empty((m, 3, 3), dtype=torch.float64, device=index.device)
    angles = torch.empty((m, 3), dtype=torch.float64, device=index.device)
    costs = torch.empty(m, dtype=torch.float64, device=index.device)
    unc = torch.empty(1, dtype=torch.int32, device=index.device)
    lib = _lib.lib()
    ws = _lib.workspace(lib.csg_rotation_refine_workspace_bytes(ns, m), index.device)
    _lib.check(lib.csg_rotation_refine(
        _lib.ptr(inputs.source_points), _lib.ptr(inputs.source_values), ns,
        _lib.ptr(inputs.target_values), _lib.ptr(index.target_faces),
        _lib.ptr(index.target_triangles), _lib.ptr(index.face_spheres()),
        _lib.ptr(index.candidate_offsets), _lib.ptr(index.candidate_faces),
        int(index.grid_size), _lib.ptr(st),
        int(start_is_angles), m, int(steps), float(max_step), float(damping), _lib.ptr(mats),
        _lib.ptr(angles), _lib.ptr(costs), _lib.ptr(ws), ws.numel(), _lib.ptr(unc),
        _lib.stream_handle(index.device)), "csg_rotation_refine")
    return mats, angles, costs, unc


def refine_adjoint_many(problems, starts, *, steps: int = 20, max_step: float = 0.05,
                        damping: float = 1e-3, start_is_angles: bool = False):
    """Gradient refinement of candidate rotations for several problems.

    ``problems`` holds (index, prepared inputs); ``starts[k]`` holds problem
    k's starting candidates (matrices, or Euler angles with
    ``start_is_angles``).  Every candidate takes ``steps`` Levenberg-Marquardt
    steps on the device (csg_rotation_refine); all problems are enqueued
    before one synchronisation.  Returns per problem (matrices [m, 3, 3],
    Euler angles [m, 3], costs [m] in the reference's angle formulation) as
    host arrays.
    """

    if steps < 0 or max_step <= 0.0 or damping <= 0.0:
        raise ValueError("steps must be non-negative; max_step and damping positive")
    launched = []
    for (index, inputs), start in zip(problems, starts):
        if not isinstance(start, torch.Tensor):
            start = torch.as_tensor(np.asarray(start, dtype=np.float64))
        launched.append(refine_adjoint_launch(index, inputs, start,
                                              start_is_angles=start_is_angles, steps=steps,
                                              max_step=max_step, damping=damping))
    host = torch.cat([torch.cat([m.reshape(-1), a.reshape(-1), c, u.to(torch.float64)])
                      for m, a, c, u in launched]).cpu().numpy()
    out, pos = [], 0
    for m, _, _, _ in launched:
        k = int(m.shape[0])
        mats = host[pos:pos + 9 * k].reshape(k, 3, 3)
        angles = host[pos + 9 * k:pos + 12 * k].reshape(k, 3)
        costs = host[pos + 12 * k:pos + 13 * k]
        if host[pos + 13 * k]:
            raise RuntimeError("Rotation is not covered by target sphere candidate index")
        pos += 13 * k + 1
        out.append((mats, angles, costs))
    return out


_HOPF_DEVICE: dict[str, torch.Tensor] = {}


@functools.lru_cache(maxsize=1)
def _hopf_host() -> np.ndarray:
    return hopf_grid()


def _top_k(costs: torch.Tensor, k: int) -> torch.Tensor:
    """Indices (int64, device) of the k first costs in stable ascending order
    (csg_top_k: ties by index, NaN last, as torch.sort(stable=True))."""

    c = costs.to(torch.float64).contiguous()
    k = min(int(k), int(c.numel()))
    order = torch.empty(k, dtype=torch.int32, device=c.device)
    _lib.check(_lib.lib().csg_top_k(_lib.ptr(c), int(c.numel()), k, _lib.ptr(order),
                                    _lib.stream_handle(c.device)), "csg_top_k")
    return order.to(torch.int64)


_SEEDS_DEVICE: dict[str, torch.Tensor] = {}


def _seeds_device(device) -> torch.Tensor:
    """The 46 seed angle triples as a [46, 3] float64 device tensor (cached, so
    that a captured schedule copies nothing from the host)."""

    key = str(device)
    if key not in _SEEDS_DEVICE:
        _SEEDS_DEVICE[key] = torch.as_tensor(official_seed_grid_angles(), dtype=torch.float64,
                                             device=device)
    return _SEEDS_DEVICE[key]


def _hopf_device(device) -> torch.Tensor:
    """The 8,192-rotation Hopf grid as a [8192, 9] float64 device tensor (cached)."""

    key = str(device)
    if key not in _HOPF_DEVICE:
        _HOPF_DEVICE[key] = torch.as_tensor(_hopf_host().reshape(-1, 9), dtype=torch.float64,
                                            device=device)
    return _HOPF_DEVICE[key]


def search_rotations(problems, *, refine: bool = True, max_iter: int = 500, tol: float = 1.0e-4,
                     simplex_step: float = 0.1, method: str = "nelder-mead",
                     candidates: str = "seeds", top: int = 8,
                     steps: int = 20) -> list[RotationSearchResult]:
    """search_rotation_prepared for several independent problems ((index,
    prepared inputs) pairs): the 46-seed scans are launched together and read
    with one synchronisation, then the Nelder-Mead refinements run in lockstep
    (refine_nelder_mead_many).  Each result equals the single-problem search.

    ``method="adjoint"`` replaces Nelder-Mead: the ``top`` best candidates
    (of the 46 seeds, or of the 8,192-rotation Hopf grid with
    ``candidates="hopf"``) take ``steps`` gradient (Levenberg-Marquardt) steps
    each (refine_adjoint_many), and the result is the best refined rotation,
    expressed as Euler angles and costed in the reference's angle formulation.
    """

    if method not in {"nelder-mead", "adjoint"}:
        raise ValueError("rotation search method must be 'nelder-mead' or 'adjoint'")
    if candidates not in {"seeds", "hopf"}:
        raise ValueError("rotation candidates must be 'seeds' or 'hopf'")
    if method == "adjoint":
        return _search_rotations_adjoint(problems, refine=refine, candidates=candidates,
                                         top=top, steps=steps)
    seeds_np = official_seed_grid_angles()
    launched = []
    for index, inputs in problems:
        seeds = torch.as_tensor(seeds_np, dtype=torch.float64, device=index.device)
        unc = torch.zeros(1, dtype=torch.int32, device=index.device)
        costs, _, _, _ = index._launch_costs(inputs, seeds, uncovered_dev=unc)
        launched.append((seeds, costs, unc))
    uncovered = torch.cat([u for _, _, u in launched]).cpu()
    if bool(uncovered.any()):
        raise RuntimeError("Rotation is not covered by target sphere candidate index")
    firsts = []
    for seeds, costs, _ in launched:
        best = int(torch.argmin(costs).item())
        firsts.append((seeds[best], costs[best], costs))
    if refine:
        refined = refine_nelder_mead_many(
            [(index, inputs, first[0]) for (index, inputs), first in zip(problems, firsts)],
            max_iter=max_iter, tol=tol, simplex_step=simplex_step)
    else:
        refined = [(a.clone(), c.clone(), 0) for a, c, _ in firsts]
    out = []
    for (index, _), (seed_angle, seed_cost, seed_costs), (angle, cost, iterations) in zip(
            problems, firsts, refined):
        out.append(RotationSearchResult(seed_angle=seed_angle.clone(), seed_cost=seed_cost.clone(),
                                        seed_costs=seed_costs, angle=angle, cost=cost,
                                        rotation_matrix=index.rotation_matrix(angle),
                                        iterations=iterations))
    return out


def _adjoint_results_on_device(problems, staged, candidates: str, refine: bool,
                               steps: int) -> list[RotationSearchResult]:
    """_search_rotations_adjoint's selection without a host read: the best
    refined candidate (first minimum, as numpy's argmin) picked on the device,
    its angles and cost kept there; the coverage counters become deferred
    checks.  seed_angle is the pool's first candidate in cost order (device)."""

    out = []
    for problem, (costs, order, angles, final, unc, unc2) in zip(problems, staged):
        index = problem[0]
        dev = index.device
        best = torch.argmin(final).reshape(1)
        angle = angles.index_select(0, best).reshape(3)
        cost = final.index_select(0, best).reshape(())
        schedule.defer(unc, "rotation not covered by the target sphere's candidate index")
        schedule.defer(unc2, "rotation not covered by the target sphere's candidate index")
        first = order[:1]
        if candidates == "seeds":
            seed_angle = _seeds_device(dev).index_select(0, first).reshape(3)
        else:
            seed_angle = torch.full((3,), float("nan"), dtype=torch.float64, device=dev)
        out.append(RotationSearchResult(
            seed_angle=seed_angle, seed_cost=costs.index_select(0, first).reshape(()),
            seed_costs=costs, angle=angle, cost=cost, rotation_matrix=None,
            iterations=steps if refine else 0))
    return out


def map_costs(source_map: torch.Tensor, target_map: torch.Tensor, matrices: torch.Tensor, *,
              stride: int = 4) -> torch.Tensor:
    """Area-weighted SSD of candidate rotations on two [ny, nx] feature maps
    (csg_rotation_map_costs): a coarse score without point location."""

    src = source_map.to(torch.float64).contiguous()
    tgt = target_map.to(torch.float64).contiguous()
    if src.ndim != 2 or tuple(src.shape) != tuple(tgt.shape):
        raise ValueError("feature maps must be two [ny, nx] arrays of one shape")
    ny, nx = int(src.shape[0]), int(src.shape[1])
    mats = matrices.to(device=src.device, dtype=torch.float32).reshape(-1, 9).contiguous()
    na = int(mats.shape[0])
    out = torch.empty(na, dtype=torch.float64, device=src.device)
    lib = _lib.lib()
    ws = _lib.workspace(lib.csg_rotation_map_costs_workspace_bytes(nx, ny, stride, na),
                        src.device)
    _lib.check(lib.csg_rotation_map_costs(_lib.ptr(src), _lib.ptr(tgt), nx, ny, int(stride),
                                          _lib.ptr(mats), na, _lib.ptr(out), _lib.ptr(ws),
                                          ws.numel(), _lib.stream_handle(src.device)),
               "csg_rotation_map_costs")
    return out


def _search_rotations_adjoint(problems, *, refine: bool, candidates: str, top: int,
                              steps: int) -> list[RotationSearchResult]:
    """Candidate scan, top-k and gradient refinement, enqueued for every problem
    and read back with one synchronisation.  A problem given with feature maps
    ((index, inputs, (source map, target map))) has the Hopf grid scored on the
    maps; otherwise every candidate is costed by point location."""

    if not 1 <= top <= 64:
        raise ValueError("top must be between 1 and 64")
    staged = []
    for problem in problems:
        index, inputs = problem[0], problem[1]
        maps = problem[2] if len(problem) > 2 else None
        dev = index.device
        unc = torch.zeros(1, dtype=torch.int32, device=dev)
        if candidates == "seeds":
            pool = _seeds_device(dev)
            costs, _, _, _ = index._launch_costs(inputs, pool, uncovered_dev=unc)
        elif maps is not None:
            pool = _hopf_device(dev)
            costs = map_costs(maps[0], maps[1], pool)
        else:
            pool = _hopf_device(dev)
            costs = index._launch_matrix_costs(inputs, pool, False, unc)
        order = _top_k(costs, top)
        start = pool.index_select(0, order)
        if refine and steps > 0:
            _, angles, final, unc2 = refine_adjoint_launch(
                index, inputs, start, start_is_angles=candidates == "seeds", steps=steps)
        else:
            _, angles, final, unc2 = refine_adjoint_launch(
                index, inputs, start, start_is_angles=candidates == "seeds", steps=0)
        staged.append((costs, order, angles, final, unc, unc2))
    if schedule.capturing():
        return _adjoint_results_on_device(problems, staged, candidates, refine, steps)
    host = torch.cat([torch.cat([order[:1].to(torch.float64), angles.reshape(-1), final,
                                 unc.to(torch.float64), unc2.to(torch.float64)])
                      for _, order, angles, final, unc, unc2 in staged]).cpu().numpy()
    out, pos = [], 0
    for problem, (costs, order, angles, final, _, _) in zip(problems, staged):
        index = problem[0]
        k = int(final.numel())
        first = int(host[pos])
        ang = host[pos + 1:pos + 1 + 3 * k].reshape(k, 3)
        vals = host[pos + 1 + 3 * k:pos + 1 + 4 * k]
        if host[pos + 1 + 4 * k] or host[pos + 2 + 4 * k]:
            raise RuntimeError("Rotation is not covered by target sphere candidate index")
        pos += 3 + 4 * k
        dev = index.device
        if candidates == "seeds":
            seed_angle = official_seed_grid_angles()[first]
        else:
            seed_angle = angles_from_matrix(_hopf_host()[first])
        best = int(np.argmin(vals))
        angle = torch.as_tensor(ang[best], dtype=torch.float64, device=dev)
        out.append(RotationSearchResult(
            seed_angle=torch.as_tensor(seed_angle, dtype=torch.float64, device=dev),
            seed_cost=costs[first], seed_costs=costs, angle=angle,
            cost=torch.as_tensor(vals[best], dtype=torch.float64, device=dev),
            rotation_matrix=index.rotation_matrix(angle),
            iterations=steps if refine else 0))
    return out
