"""The fused multigrid schedule computes exactly what the per-operator one does.

The fused schedule (fine levels relaxed on LDS tiles with temporal blocking,
coarse levels in one workgroup) regroups the reference's operator calls
(dartel_grid.py:1211-1417) without changing any pixel's arithmetic or its
order, so its results must be bit-identical to the one-launch-per-operator
schedule, which is itself pinned to the reference golden vectors and the
oracle.  Shapes cover: the production 512x256 pyramid (three tiled levels),
a width that is not a multiple of the tile width with fewer rows than one tile
(1000x8), odd widths inside the coarse workgroup (130x66, 9x7, 6x1) and a
problem that fits in one workgroup entirely (96x40).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SHAPES = [(512, 256), (1000, 8), (130, 66), (96, 40), (9, 7), (6, 1)]


@pytest.fixture(scope="module")
def csg():
    import cat_surface_gpu

    return cat_surface_gpu


def _system(nx, ny, rng, dev, dtype):
    a = np.stack([rng.uniform(0.5, 2.0, (ny, nx)), rng.uniform(0.5, 2.0, (ny, nx)),
                  rng.uniform(-0.3, 0.3, (ny, nx))])
    b = rng.standard_normal((2, ny, nx))
    u0 = 0.1 * rng.standard_normal((2, ny, nx))
    t = lambda x: torch.as_tensor(x, dtype=dtype, device=dev)  # noqa: E731
    return t(a), t(b), t(u0)


def _both(csg, fn):
    previous = csg.set_fmg_schedule(False)
    try:
        ref = fn()
        csg.set_fmg_schedule(True)
        fused = fn()
    finally:
        csg.set_fmg_schedule(previous)
    return ref, fused


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("mode", ["coarse_init", "initial", "assume_nonzero"])
def test_fused_fmg_is_bit_identical(csg, cuda, dtype, shape, mode):
    nx, ny = shape
    print(f"fmg schedule check {shape} {dtype} {mode}", flush=True)
    rng = np.random.default_rng(nx * 1009 + ny)
    a, b, u0 = _system(nx, ny, rng, cuda, dtype)
    params = (1.0, 1.0, 0.125, 0.0, 1e-3)
    kwargs = dict(cycles=3, nit=3, device=cuda, dtype=dtype)
    if mode == "initial":
        kwargs["initial"] = u0
    elif mode == "assume_nonzero":
        kwargs["initial"] = u0
        kwargs["assume_initial_nonzero"] = True
    ref, fused = _both(csg, lambda: csg.fmg2_membrane(a, b, params, **kwargs))
    assert torch.isfinite(ref).all()
    assert torch.equal(ref, fused)


@pytest.mark.parametrize("nit,cycles", [(1, 2), (2, 1), (0, 2)])
def test_fused_fmg_other_sweep_counts(csg, cuda, nit, cycles):
    rng = np.random.default_rng(nit * 10 + cycles)
    a, b, u0 = _system(256, 128, rng, cuda, torch.float64)
    params = (1.0, 1.0, 0.05, 0.0, 1e-3)
    ref, fused = _both(csg, lambda: csg.fmg2_membrane(a, b, params, cycles=cycles, nit=nit,
                                                       initial=u0, device=cuda))
    assert torch.equal(ref, fused)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_fused_schedule_solve_is_bit_identical(csg, cuda, dtype):
    """A whole DARTEL solve (every FMG of every Gauss-Newton step) at 256x128."""

    nx, ny = 256, 128
    y, x = np.mgrid[0:ny, 0:nx]
    src = np.sin(2 * np.pi * x / nx) * np.cos(np.pi * (y + 0.5) / ny)
    tgt = np.sin(2 * np.pi * (x - 3.0) / nx) * np.cos(np.pi * (y + 1.5) / ny)
    s = torch.as_tensor(src[None], dtype=dtype, device=cuda)
    t = torch.as_tensor(tgt[None], dtype=dtype, device=cuda)
    ref, fused = _both(csg, lambda: csg.solve_dartel_maps(s, t, loop=3, device=cuda,
                                                          dtype=dtype).flow)
    assert torch.equal(ref, fused)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("tile", [(64, 8), (16, 16), (128, 4), (2, 1), (34, 5)])
def test_relaxation_tile_shape_does_not_change_values(csg, cuda, dtype, tile):
    """Every fine level relaxed on tiles of another shape gives the same bits."""

    from cat_surface_gpu.autotune import fine_levels, set_tile

    nx, ny = 512, 256
    rng = np.random.default_rng(77)
    a, b, u0 = _system(nx, ny, rng, cuda, dtype)
    params = (1.0, 1.0, 0.125, 0.0, 1e-3)
    ref = csg.fmg2_membrane(a, b, params, device=cuda, dtype=dtype)
    levels = fine_levels(nx, ny, dtype)
    assert levels and levels[0] == (nx, ny)
    try:
        for lx, ly in levels:
            set_tile(lx, ly, dtype, tile)
        other = csg.fmg2_membrane(a, b, params, device=cuda, dtype=dtype)
        with_initial = csg.fmg2_membrane(a, b, params, initial=u0, device=cuda, dtype=dtype)
    finally:
        for lx, ly in levels:
            set_tile(lx, ly, dtype, None)
    assert torch.equal(ref, other)
    assert torch.equal(with_initial, csg.fmg2_membrane(a, b, params, initial=u0, device=cuda,
                                                       dtype=dtype))


def test_relaxation_tile_rejects_bad_shapes(csg, cuda):
    from cat_surface_gpu.autotune import set_tile

    for bad in ((3, 8), (0, 4), (4, 0), (512, 512)):
        with pytest.raises(ValueError):
            set_tile(512, 256, torch.float32, bad)


def test_autotuner_installs_valid_shapes(csg, cuda, tmp_path, monkeypatch):
    from cat_surface_gpu import autotune

    monkeypatch.setenv("CSG_AUTOTUNE_CACHE", str(tmp_path / "tiles.json"))
    autotune._TUNED.clear()
    try:
        chosen = autotune.ensure_tuned([(128, 64)], torch.float32, 4, device=cuda)
        levels = chosen[(128, 64)]
        assert set(levels) == set(autotune.fine_levels(128, 64, torch.float32))
        for (lx, ly), shape in levels.items():
            assert autotune.current_tile(lx, ly, torch.float32) == tuple(shape)
        assert (tmp_path / "tiles.json").is_file()
        # a second process-level call reads the cache instead of tuning
        autotune._TUNED.clear()
        again = autotune.ensure_tuned([(128, 64)], torch.float32, 4, device=cuda)
        assert {k: tuple(v) for k, v in again[(128, 64)].items()} == \
            {k: tuple(v) for k, v in levels.items()}
    finally:
        for lx, ly in autotune.fine_levels(128, 64, torch.float32):
            autotune.set_tile(lx, ly, torch.float32, None)
        autotune._TUNED.clear()
