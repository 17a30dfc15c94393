"""End-to-end parity of the HIP CAT_SurfWarp pipeline with the oracle chain.

The oracle chain (oracle/pipeline.py) composes the oracle's pinned stages into
the whole registration of src/cat_surface_gpu/gpu_pipeline.py:468-1326 on the
CPU, from the same four input meshes: its own stencils, rotation features,
46-seed scan and Nelder-Mead, DARTEL runs (float64) and warps.  The HIP
pipeline's output sphere is compared with the chain's the way the reference
compares meshes (tools/benchmark_cat_surface_end_to_end.py:79-103): faces
exactly equal, and the max / mean / p99 vertex error.

Workloads (synthetic, seeded):
- icosphere-6 pair (40,962 vertices), SH degree <= 8, target rotated <= 15
  degrees and warped <= 3 degrees; steps=1 runs=1, no avg, 512x256 map;
- 61,442-vertex jittered-Fibonacci-hull hemisphere pair, SH degree <= 16 +
  N(0, 0.05) noise, target rotated <= 30 degrees and warped; steps=2 runs=2
  avg, left and x-mirrored right copies.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SEED = 20261203

WORKLOADS = {
    "icosphere6": dict(pair=dict(kind="icosphere", subdivisions=6, lmax=8, noise=0.0,
                                 max_rotation_deg=15.0, warp_deg=3.0, seed=SEED + 1),
                       run=dict(steps=1, runs=1, avg=False)),
    "fibonacci61442-lh": dict(pair=dict(kind="fibonacci", n=61442, lmax=16, noise=0.05,
                                        max_rotation_deg=30.0, warp_deg=3.0, seed=SEED + 2),
                              run=dict(steps=2, runs=2, avg=True)),
    "fibonacci61442-rh": dict(pair=dict(kind="fibonacci", n=61442, lmax=16, noise=0.05,
                                        max_rotation_deg=30.0, warp_deg=3.0, seed=SEED + 2,
                                        mirror=True),
                              run=dict(steps=2, runs=2, avg=True)),
}

_CACHE: dict = {}


def _pair(name):
    from cat_surface_gpu import synthetic

    key = ("pair", name)
    if key not in _CACHE:
        _CACHE[key] = synthetic.make_pair(**WORKLOADS[name]["pair"])
    return _CACHE[key]


def _oracle(name):
    from oracle import pipeline as P

    key = ("oracle", name)
    if key not in _CACHE:
        p = _pair(name)
        print(f"\n[{name}] oracle chain ...", flush=True)
        res, wall, cpu = P.timed_surfwarp(p.src_surface, p.src_sphere, p.src_faces,
                                          p.tgt_surface, p.tgt_sphere, p.tgt_faces,
                                          **WORKLOADS[name]["run"])
        print(f"[{name}] oracle chain: {wall:.1f} s wall, {cpu:.1f} s CPU, rotation "
              f"{np.round(res.rotation.angle, 6)} cost {res.rotation.cost:.6f}", flush=True)
        _CACHE[key] = res
    return _CACHE[key]


def _gpu(name, dtype, cuda):
    from cat_surface_gpu import GiftiMesh, run_cat_surface_gpu_pipeline

    p = _pair(name)
    meshes = (GiftiMesh(p.src_surface, p.src_faces), GiftiMesh(p.src_sphere, p.src_faces),
              GiftiMesh(p.tgt_surface, p.tgt_faces), GiftiMesh(p.tgt_sphere, p.tgt_faces))
    return run_cat_surface_gpu_pipeline(None, None, None, None, None, meshes=meshes,
                                        write_output=False, device=cuda, dartel_dtype=dtype,
                                        **WORKLOADS[name]["run"])


def _errors(a, b):
    d = np.linalg.norm(np.asarray(a, np.float64) - np.asarray(b, np.float64), axis=1)
    return float(d.max()), float(d.mean()), float(np.percentile(d, 99))


# FP64 output spheres are float32 arrays: a few float32 ulps (6e-8 at unit radius)
# apart at best; the bound also covers float32 rounding of the intermediate spheres.
TOLERANCE = {"float32": 1e-3, "float64": 1e-5}


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("name", list(WORKLOADS))
def test_surfwarp_matches_oracle_chain(name, dtype, cuda):
    ref = _oracle(name)
    p = _pair(name)
    res = _gpu(name, dtype, cuda)
    np.testing.assert_array_equal(res.faces, p.src_faces)
    assert res.vertices.dtype == np.float32 and res.vertices.shape == ref.vertices.shape
    mx, mean, p99 = _errors(res.vertices, ref.vertices)
    dang = float(np.abs(np.asarray(res.rotation_angle) - ref.rotation.angle).max())
    print(f"[{name}] {dtype}: vertex error max {mx:.3e} mean {mean:.3e} p99 {p99:.3e}; "
          f"rotation |d angle| {dang:.2e} rad, cost {res.rotation_cost:.6f} vs "
          f"{ref.rotation.cost:.6f}", flush=True)
    # the rotation is the same Nelder-Mead trajectory up to cost rounding
    assert dang <= 1e-4
    assert mx <= TOLERANCE[dtype]


def test_per_step_velocities_float32_and_oracle_512x256(cuda):
    """Every Gauss-Newton step of both 512x256 solves of the 61,442-vertex pair's
    first DARTEL run: the float32 velocity within relative L2 1e-4 of the float64
    one, and the float64 one within 1e-9 of the oracle's, each trajectory
    advancing from its own previous velocity."""

    import cat_surface_gpu as csg
    from oracle import oracle as O
    from oracle import pipeline as P

    p = _pair("fibonacci61442-lh")
    src_st = P.build_stencil(p.src_sphere, p.src_faces, normalise=True)
    tgt_st = P.build_stencil(p.tgt_sphere, p.tgt_faces, normalise=True)
    fwhm = (5.0 / 9.0, 5.0 / 27.0)
    src = P.type5_maps(p.src_surface, src_st, fwhm)
    tgt = P.type5_maps(p.tgt_surface, tgt_st, fwhm)
    params = O.default_dartel_parameters(6)
    worst32 = worst64 = 0.0
    for s in range(2):
        maps = {dt: (torch.as_tensor(src[s], dtype=dt, device=cuda),
                     torch.as_tensor(tgt[s], dtype=dt, device=cuda))
                for dt in (torch.float32, torch.float64)}
        v = {dt: torch.zeros((2, 256, 512), dtype=dt, device=cuda)
             for dt in (torch.float32, torch.float64)}
        vo = np.zeros((2, 256, 512))
        for k in range(6):
            for _ in range(3):
                for dt in (torch.float32, torch.float64):
                    v[dt], _ = csg.dartel_step(*maps[dt], v[dt], k=k, params=tuple(params[k]),
                                               device=cuda, dtype=dt)
                vo, _ = O.dartel_step(src[s], tgt[s], vo, k, tuple(params[k]))
                v64 = v[torch.float64].cpu().numpy()
                v32 = v[torch.float32].double().cpu().numpy()
                scale = np.linalg.norm(v64)
                worst32 = max(worst32, float(np.linalg.norm(v32 - v64) / scale))
                worst64 = max(worst64, float(np.linalg.norm(v64 - vo) / scale))
    print(f"\nper-step velocity: float32 vs float64 worst rel L2 {worst32:.2e}; float64 vs "
          f"oracle {worst64:.2e}", flush=True)
    assert worst32 <= 1e-4
    assert worst64 <= 1e-9
