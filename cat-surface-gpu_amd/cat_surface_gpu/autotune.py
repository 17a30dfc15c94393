"""Autotuner for the fused multigrid's fine-level relaxation tiles.

The fused FMG schedule relaxes every fine level (those above the single-
workgroup coarse levels) on LDS tiles with a halo.  Any tile shape (even
width) gives bit-identical values, so the shape is a pure speed knob; the
best one depends on the level size, the batch (independent solves in one
launch) and the dtype: small maps want small tiles to fill the 256 CUs,
large maps want wide tiles to cut halo re-reads.

``tune_relax_tiles`` times whole FMG solves (csg_fmg2_membrane) on synthetic
data while it walks the fine levels one at a time, trying each candidate
shape for that level with the others fixed, and installs the fastest
(csg_set_relax_tile).  ``ensure_tuned`` does this once per (map size, dtype,
batch) per process and keeps the choices in a JSON cache
(``CSG_AUTOTUNE_CACHE``, default ~/.cache/cat_surface_gpu/relax_tiles.json)
so later processes only install them.  Tune before capturing a graph: a
captured graph keeps the shapes it was captured with.
"""

from __future__ import annotations

import ctypes
import json
import os
from pathlib import Path
from typing import Iterable, Sequence

import torch

from . import _lib

CANDIDATES: tuple[tuple[int, int], ...] = (
    (32, 16), (64, 8), (16, 16), (32, 8), (64, 16), (16, 32), (128, 4), (32, 32), (16, 8),
    (64, 32), (128, 16), (128, 8))

_TUNED: dict[tuple, dict[tuple[int, int], tuple[int, int]]] = {}
_SOURCE: dict[tuple, str] = {}


def fine_levels(nx: int, ny: int, dtype: torch.dtype, nit: int = 3) -> list[tuple[int, int]]:
    """The level sizes the fused FMG relaxes on tiles (csg_fmg_tiled_levels)."""

    count = _lib.lib().csg_fmg_tiled_levels(nx, ny, nit, _lib.dtype_code(dtype))
    if count < 0:
        _lib.check(count, "csg_fmg_tiled_levels")
    out = []
    for _ in range(count):
        out.append((nx, ny))
        nx, ny = (nx + 1) // 2, (ny + 1) // 2
    return out


def current_tile(nx: int, ny: int, dtype: torch.dtype) -> tuple[int, int]:
    tx, ty = ctypes.c_int(), ctypes.c_int()
    _lib.check(_lib.lib().csg_get_relax_tile(nx, ny, _lib.dtype_code(dtype), ctypes.byref(tx),
                                             ctypes.byref(ty)), "csg_get_relax_tile")
    return int(tx.value), int(ty.value)


def set_tile(nx: int, ny: int, dtype: torch.dtype, shape: tuple[int, int] | None) -> None:
    tx, ty = (0, 0) if shape is None else shape
    _lib.check(_lib.lib().csg_set_relax_tile(nx, ny, _lib.dtype_code(dtype), int(tx), int(ty)),
               "csg_set_relax_tile")


def _fmg_timer(nx: int, ny: int, batch: int, dtype: torch.dtype, device: torch.device):
    g = torch.Generator(device="cpu").manual_seed(1234)
    a = torch.rand((batch, 3, ny, nx), generator=g, dtype=torch.float64)
    a[:, 2] -= 0.5
    a[:, 2] *= 0.2
    hes = a.to(dtype=dtype, device=device).contiguous()
    rhs = (torch.rand((batch, 2, ny, nx), generator=g, dtype=torch.float64) - 0.5).to(
        dtype=dtype, device=device).contiguous()
    out = torch.empty_like(rhs)
    code = _lib.dtype_code(dtype)
    ws = _lib.workspace(_lib.lib().csg_fmg2_workspace_bytes(nx, ny, batch, code), device)
    params = _lib.doubles((1.0, 1.0, 0.125, 0.0, 1e-3))
    stream = _lib.stream_handle(device)

    def run(repeats: int) -> float:
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(repeats):
            _lib.check(_lib.lib().csg_fmg2_membrane(
                _lib.ptr(hes), _lib.ptr(rhs), params, 3, 3, None, 1, nx, ny, batch, code,
                _lib.ptr(out), _lib.ptr(ws), ws.numel(), stream), "csg_fmg2_membrane")
        stop.record()
        stop.synchronize()
        return start.elapsed_time(stop) / repeats

    return run


def tune_relax_tiles(nx: int, ny: int, dtype: torch.dtype = torch.float32, batch: int = 1, *,
                     candidates: Sequence[tuple[int, int]] = CANDIDATES, repeats: int = 5,
                     device="cuda") -> dict:
    """Tune and install the tile shape of every fine level of an nx x ny FMG;
    returns {"levels": {(lnx, lny): (tx, ty)}, "ms": best FMG time, "default_ms": ...}."""

    dev = torch.device(device)
    levels = fine_levels(nx, ny, dtype)
    for lx, ly in levels:
        set_tile(lx, ly, dtype, None)
    run = _fmg_timer(nx, ny, batch, dtype, dev)
    run(2)
    default_ms = min(run(repeats) for _ in range(3))
    best_ms = default_ms
    chosen: dict[tuple[int, int], tuple[int, int]] = {}
    for lx, ly in levels:
        best_shape = None
        for shape in candidates:
            if shape == (32, 16):
                continue  # the default, already timed
            try:
                set_tile(lx, ly, dtype, shape)
            except (ValueError, RuntimeError):
                continue  # does not fit the LDS budget
            run(1)
            ms = min(run(repeats) for _ in range(3))
            if ms < best_ms * 0.99:
                best_ms, best_shape = ms, shape
        if best_shape is None:
            set_tile(lx, ly, dtype, None)
            chosen[(lx, ly)] = (32, 16)
        else:
            set_tile(lx, ly, dtype, best_shape)
            chosen[(lx, ly)] = best_shape
    return {"levels": chosen, "ms": best_ms, "default_ms": default_ms}


def _cache_path() -> Path:
    env = os.environ.get("CSG_AUTOTUNE_CACHE")
    if env:
        return Path(env)
    return Path.home() / ".cache" / "cat_surface_gpu" / "relax_tiles.json"


def _load_cache() -> dict:
    try:
        return json.loads(_cache_path().read_text())
    except (OSError, ValueError):
        return {}


def _store_cache(data: dict) -> None:
    try:
        path = _cache_path()
        path.parent.mkdir(parents=True, exist_ok=True)
        path.write_text(json.dumps(data, indent=1, sort_keys=True))
    except OSError:
        pass  # a read-only home only costs a re-tune in the next process


def ensure_tuned(shapes: Iterable[tuple[int, int]], dtype: torch.dtype, batch: int, *,
                 device="cuda", use_cache: bool = True) -> dict:
    """Tile shapes for every (nx, ny) map size, tuned at most once per process."""

    dev = torch.device(device)
    props = torch.cuda.get_device_properties(dev)
    name = f"{getattr(props, 'gcnArchName', props.name)}:{props.multi_processor_count}cu"
    cache = _load_cache() if use_cache else {}
    changed = False
    out = {}
    for nx, ny in sorted(set((int(a), int(b)) for a, b in shapes)):
        key = (name, nx, ny, str(dtype), int(batch))
        if key not in _TUNED:
            ckey = f"{name}|{nx}x{ny}|{str(dtype).replace('torch.', '')}|b{int(batch)}"
            entry = cache.get(ckey)
            if entry is not None:
                levels = {tuple(int(v) for v in k.split("x")): tuple(s)
                          for k, s in entry["levels"].items()}
                for (lx, ly), shape in levels.items():
                    set_tile(lx, ly, dtype, shape)
                _SOURCE[key] = "cache"
            else:
                res = tune_relax_tiles(nx, ny, dtype, batch, device=dev)
                levels = res["levels"]
                cache[ckey] = {"levels": {f"{a}x{b}": list(s) for (a, b), s in levels.items()},
                               "ms": res["ms"], "default_ms": res["default_ms"]}
                changed = True
                _SOURCE[key] = "tuned"
            _TUNED[key] = levels
        out[(nx, ny)] = _TUNED[key]
    if changed and use_cache:
        _store_cache(cache)
    return out


def installed() -> list[dict]:
    """The tile shapes ensure_tuned installed in this process, with their source."""

    return [{"map": f"{k[1]}x{k[2]}", "dtype": k[3].replace("torch.", ""), "batch": k[4],
             "source": _SOURCE.get(k, "?"),
             "levels": {f"{a}x{b}": list(v) for (a, b), v in levels.items()}}
            for k, levels in _TUNED.items()]
