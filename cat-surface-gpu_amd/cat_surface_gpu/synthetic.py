"""Seeded synthetic hemisphere pairs for tests and benchmarks.

Two mesh families:
- icosphere of subdivision k (10·4^k + 2 vertices);
- convex hull of a jittered Fibonacci sphere (jitter as a fraction of the
  point spacing), outward-wound.
Features are real spherical-harmonic sums of degree <= lmax with
N(0, 1/l^1.5) coefficients (plus optional N(0, sigma) noise).  The target
carries the source feature rotated by a seeded rotation and displaced by a
smooth tangent field.  White surfaces are sphere · (1 + 0.05 f) · radius_mm:
heat smoothing runs ⌈fwhm²·0.541011/ē²⌉ iterations with fwhm 50 fixed
(src/cat_surface_gpu/surface_stencil.py:365, :511), which stays a few hundred
only at millimetre scale; the spheres themselves are unit radius.
"""

from __future__ import annotations

from dataclasses import dataclass
import math

import numpy as np
from scipy.spatial import ConvexHull
from scipy.spatial.transform import Rotation

from .stencil_builder import icosphere


@dataclass(frozen=True)
class SyntheticPair:
    src_sphere: np.ndarray
    src_surface: np.ndarray
    src_faces: np.ndarray
    tgt_sphere: np.ndarray
    tgt_surface: np.ndarray
    tgt_faces: np.ndarray
    rotation: np.ndarray  # 3x3 applied to the source feature


def fibonacci_hull(n: int, jitter: float, seed: int) -> tuple[np.ndarray, np.ndarray]:
    rng = np.random.default_rng(seed)
    i = np.arange(n, dtype=np.float64) + 0.5
    z = 1.0 - 2.0 * i / n
    phi = math.pi * (1.0 + 5.0 ** 0.5) * i
    r = np.sqrt(1.0 - z * z)
    pts = np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)
    spacing = math.sqrt(4.0 * math.pi / n)
    pts = pts + rng.normal(size=pts.shape) * (jitter * spacing / math.sqrt(3.0))
    pts /= np.linalg.norm(pts, axis=1, keepdims=True)
    hull = ConvexHull(pts)
    faces = hull.simplices.astype(np.int64)
    centres = pts[faces].mean(axis=1)
    normal = np.cross(pts[faces[:, 1]] - pts[faces[:, 0]], pts[faces[:, 2]] - pts[faces[:, 0]])
    flip = (normal * centres).sum(1) < 0
    faces[flip] = faces[flip][:, [0, 2, 1]]
    if np.unique(faces).size != n:
        raise RuntimeError("convex hull dropped vertices")
    return pts, faces.astype(np.int32)


def real_sh_basis(points: np.ndarray, lmax: int) -> np.ndarray:
    """[n, (lmax+1)^2 - 1] orthonormal real spherical harmonics of degree 1..lmax,
    columns l = 1..lmax, m = -l..l: sqrt(2) P_l^|m|(cos t) sin(|m| p) for m < 0,
    P_l^0(cos t) for m = 0 and sqrt(2) P_l^m(cos t) cos(m p) for m > 0, with the
    fully normalised associated Legendre functions (no Condon-Shortley phase)
    from the standard three-term recurrences."""

    x, y, z = points[:, 0], points[:, 1], points[:, 2]
    ct = np.clip(z, -1.0, 1.0)
    st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
    phi = np.arctan2(y, x)
    # pbar[l][m] for 0 <= m <= l <= lmax
    pbar = [[None] * (l + 1) for l in range(lmax + 1)]
    pbar[0][0] = np.full_like(ct, math.sqrt(1.0 / (4.0 * math.pi)))
    for m in range(1, lmax + 1):
        pbar[m][m] = math.sqrt((2.0 * m + 1.0) / (2.0 * m)) * st * pbar[m - 1][m - 1]
    for m in range(0, lmax):
        pbar[m + 1][m] = math.sqrt(2.0 * m + 3.0) * ct * pbar[m][m]
    for m in range(0, lmax + 1):
        for l in range(m + 2, lmax + 1):
            a = math.sqrt((4.0 * l * l - 1.0) / (l * l - m * m))
            b = math.sqrt(((l - 1.0) ** 2 - m * m) / (4.0 * (l - 1.0) ** 2 - 1.0))
            pbar[l][m] = a * (ct * pbar[l - 1][m] - b * pbar[l - 2][m])
    cos_m = [np.cos(m * phi) for m in range(lmax + 1)]
    sin_m = [np.sin(m * phi) for m in range(lmax + 1)]
    cols = []
    root2 = math.sqrt(2.0)
    for l in range(1, lmax + 1):
        for m in range(-l, l + 1):
            if m < 0:
                cols.append(root2 * pbar[l][-m] * sin_m[-m])
            elif m == 0:
                cols.append(pbar[l][0])
            else:
                cols.append(root2 * pbar[l][m] * cos_m[m])
    return np.stack(cols, axis=1)


def sh_feature(points: np.ndarray, coeffs: np.ndarray, lmax: int) -> np.ndarray:
    f = real_sh_basis(points, lmax) @ coeffs
    return (f - f.mean()) / f.std()


def sh_coefficients(lmax: int, rng) -> np.ndarray:
    out = []
    for l in range(1, lmax + 1):
        out.append(rng.normal(scale=math.sqrt(1.0 / l ** 1.5), size=2 * l + 1))
    return np.concatenate(out)


def tangent_warp(points: np.ndarray, amplitude_deg: float, rng) -> np.ndarray:
    """Smooth tangent displacement (a few low-order gradient fields), normalised back."""

    coeffs = rng.normal(size=(3, 8))
    p = points
    field = np.zeros_like(p)
    for k in range(3):
        a, b, c, d, e, f, g, h = coeffs[k]
        grad = np.stack([a + d * p[:, 1] + g * p[:, 2], b + d * p[:, 0] + e * p[:, 2],
                         c + e * p[:, 1] + g * p[:, 0] + h * p[:, 2]], axis=1)
        field += grad
    field -= (field * p).sum(1, keepdims=True) * p
    scale = math.radians(amplitude_deg) / max(np.abs(field).max(), 1e-12)
    out = p + field * scale
    return out / np.linalg.norm(out, axis=1, keepdims=True)


def make_pair(kind: str = "fibonacci", n: int = 61442, *, lmax: int = 16, noise: float = 0.05,
              max_rotation_deg: float = 30.0, warp_deg: float = 3.0, radius_mm: float = 100.0,
              seed: int = 20261203, subdivisions: int = 6, mirror: bool = False) -> SyntheticPair:
    rng = np.random.default_rng(seed)
    if kind == "icosphere":
        sv, sf = icosphere(subdivisions)
        tv, tf = sv.copy(), sf.copy()
    else:
        sv, sf = fibonacci_hull(n, 0.2, seed + 1)
        tv, tf = fibonacci_hull(n, 0.2, seed + 2)
    coeffs = sh_coefficients(lmax, rng)
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    angle = math.radians(max_rotation_deg) * rng.uniform(0.3, 1.0)
    rot = Rotation.from_rotvec(axis * angle).as_matrix()
    f_src = sh_feature(sv, coeffs, lmax)
    if noise > 0:
        f_src = f_src + rng.normal(scale=noise, size=f_src.shape)
    warped = tangent_warp(tv, warp_deg, rng)
    f_tgt = sh_feature(warped @ rot, coeffs, lmax)
    if noise > 0:
        f_tgt = f_tgt + rng.normal(scale=noise, size=f_tgt.shape)

    def surface(sphere, feat):
        return (sphere * (1.0 + 0.05 * np.tanh(feat))[:, None] * radius_mm).astype(np.float32)

    ss, ts = surface(sv, f_src), surface(tv, f_tgt)
    sv32, tv32 = sv.astype(np.float32), tv.astype(np.float32)
    if mirror:
        flip = np.array([-1.0, 1.0, 1.0], dtype=np.float32)
        sv32, tv32, ss, ts = sv32 * flip, tv32 * flip, ss * flip, ts * flip
        sf, tf = sf[:, [0, 2, 1]], tf[:, [0, 2, 1]]
    return SyntheticPair(sv32, ss, np.ascontiguousarray(sf, np.int32), tv32, ts,
                         np.ascontiguousarray(tf, np.int32), rot)


def write_pair(pair: SyntheticPair, folder, compress_level: int | None = None) -> dict[str, str]:
    from pathlib import Path

    from .gifti import write_gifti_surface

    folder = Path(folder)
    folder.mkdir(parents=True, exist_ok=True)
    paths = {}
    for name, v, f in (("src_surface", pair.src_surface, pair.src_faces),
                       ("src_sphere", pair.src_sphere, pair.src_faces),
                       ("tgt_surface", pair.tgt_surface, pair.tgt_faces),
                       ("tgt_sphere", pair.tgt_sphere, pair.tgt_faces)):
        p = folder / f"{name}.gii"
        write_gifti_surface(p, v, f, compress_level=compress_level)
        paths[name] = str(p)
    return paths
