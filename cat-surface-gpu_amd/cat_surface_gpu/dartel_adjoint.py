"""Differentiable DARTEL objective: exponentiation + bilinear warp + SSD on MI355X.

E(v) = ½ Σ (f(φ) - t)² + ½ Σ v·(L v), φ = exp(v) by K squarings, with the
reference's forward arithmetic (expdef, dartel_grid.py:1619-1666; the sample
and half sum of squares of initialise_objfun, :723-775; the rtype-1
regulariser with parameters (1, 1, membrane, 0, 0), :917-955).  The backward
pass is a hand-written adjoint in HIP (csg_exp_warp_ssd_backward) behind
``ExpWarpSSD``, a torch.autograd.Function.  It is the exact gradient of the
discrete forward; the reference's DARTEL step uses instead the push-forward
gradient of the same objective (``dartel_rhs`` below restates it with this
library's operators), a different discretisation of the continuous one.

``recompute=True`` (the default) keeps no intermediate maps between the
passes: the backward rebuilds each squaring's input from v (K(K+1)/2 extra
squarings, two maps of workspace); ``recompute=False`` stores the K + 1 maps.
"""

from __future__ import annotations

import time

import numpy as np
import torch

from . import _lib
from .dartel_grid import (expdef, initialise_objfun, jac_div_smalldef, regularization_operator,
                          smalldef_jac, squaring_update)


def _check(velocity: torch.Tensor, feature: torch.Tensor, target: torch.Tensor,
           squarings: int) -> tuple[int, int, int]:
    if velocity.device.type != "cuda":
        raise RuntimeError("exp_warp_ssd runs on the GPU (HIP); got a CPU tensor")
    if velocity.dtype not in (torch.float32, torch.float64):
        raise ValueError("exp_warp_ssd supports float32 and float64")
    if velocity.ndim != 3 or velocity.shape[0] != 2:
        raise ValueError("velocity must have shape [2, ny, nx]")
    ny, nx = int(velocity.shape[1]), int(velocity.shape[2])
    for name, t in (("feature", feature), ("target", target)):
        if tuple(t.shape) != (ny, nx):
            raise ValueError(f"{name} must have shape [{ny}, {nx}]")
        if t.dtype != velocity.dtype or t.device != velocity.device:
            raise ValueError(f"{name} must match the velocity's dtype and device")
    if not 0 <= squarings <= 30:
        raise ValueError("squarings must be in [0, 30]")
    return nx, ny, _lib.dtype_code(velocity.dtype)


def _workspace(nx, ny, squarings, code, store, device):
    return _lib.workspace(_lib.lib().csg_exp_warp_ssd_workspace_bytes(nx, ny, squarings, code,
                                                                       int(store)), device)


class ExpWarpSSD(torch.autograd.Function):
    """E(v) (float64, 0-d) with a hand-written HIP adjoint for dE/dv."""

    @staticmethod
    def forward(ctx, velocity, feature, target, squarings: int, membrane: float,
                recompute: bool):
        v = velocity.contiguous()
        f = feature.contiguous()
        t = target.contiguous()
        nx, ny, code = _check(v, f, t, squarings)
        store = not recompute
        ws = _workspace(nx, ny, squarings, code, store, v.device)
        energy = torch.empty((), dtype=torch.float64, device=v.device)
        _lib.check(_lib.lib().csg_exp_warp_ssd_forward(
            _lib.ptr(v), _lib.ptr(f), _lib.ptr(t), nx, ny, int(squarings), float(membrane), code,
            int(store), _lib.ptr(energy), None, _lib.ptr(ws), ws.numel(),
            _lib.stream_handle(v.device)), "csg_exp_warp_ssd_forward")
        ctx.save_for_backward(v, f, t)
        ctx.args = (nx, ny, int(squarings), float(membrane), code, store)
        ctx.ws = ws if store else None
        return energy

    @staticmethod
    def backward(ctx, grad_output):
        v, f, t = ctx.saved_tensors
        nx, ny, squarings, membrane, code, store = ctx.args
        ws = ctx.ws if store else _workspace(nx, ny, squarings, code, False, v.device)
        scale = grad_output.to(device=v.device, dtype=torch.float64).contiguous()
        grad = torch.empty_like(v)
        _lib.check(_lib.lib().csg_exp_warp_ssd_backward(
            _lib.ptr(v), _lib.ptr(f), _lib.ptr(t), nx, ny, squarings, membrane, code,
            int(store), _lib.ptr(scale), _lib.ptr(grad), _lib.ptr(ws), ws.numel(),
            _lib.stream_handle(v.device)), "csg_exp_warp_ssd_backward")
        return grad, None, None, None, None, None


def exp_warp_ssd(velocity: torch.Tensor, feature: torch.Tensor, target: torch.Tensor, *,
                 squarings: int = 6, membrane: float = 0.01,
                 recompute: bool = True) -> torch.Tensor:
    """E(v) as a differentiable 0-d float64 tensor."""

    return ExpWarpSSD.apply(velocity, feature, target, int(squarings), float(membrane),
                            bool(recompute))


def exp_warp_ssd_parts(velocity, feature, target, *, squarings: int = 6,
                       membrane: float = 0.01) -> tuple[float, float, float]:
    """(E, SSD half sum, membrane half sum) without autograd (synchronises)."""

    v = velocity.contiguous()
    nx, ny, code = _check(v, feature, target, squarings)
    ws = _workspace(nx, ny, squarings, code, False, v.device)
    out = torch.empty(3, dtype=torch.float64, device=v.device)
    _lib.check(_lib.lib().csg_exp_warp_ssd_forward(
        _lib.ptr(v), _lib.ptr(feature.contiguous()), _lib.ptr(target.contiguous()), nx, ny,
        int(squarings), float(membrane), code, 0, _lib.ptr(out), _lib.ptr(out[1:]),
        _lib.ptr(ws), ws.numel(), _lib.stream_handle(v.device)), "csg_exp_warp_ssd_forward")
    e, s, m = out.tolist()
    return e, s, m


def dartel_rhs(feature, target, velocity, *, squarings: int = 6,
               membrane: float = 0.01) -> torch.Tensor:
    """The reference DARTEL step's code-0 right-hand side before its solve
    (dartel_grid.py:1493-1592): rhs = push-forward gradient · 2^-k + L v, the
    DARTEL approximation of dE/dv, from this library's HIP operators."""

    v = velocity
    dtype, dev = v.dtype, v.device
    scale = 1.0 / float(2 ** squarings)
    phi, jac = expdef(v, squarings, return_jacobian=True, device=dev, dtype=dtype)
    jac = jac_div_smalldef(jac, v, scale, device=dev, dtype=dtype)
    _, rhs, hes = initialise_objfun(feature, target, phi, jac, device=dev, dtype=dtype)
    tmap, tjac = smalldef_jac(v, -scale, device=dev, dtype=dtype)
    rhs, _, _, _ = squaring_update(rhs, hes, tmap, tjac, squarings, device=dev, dtype=dtype)
    reg = regularization_operator(v, (1.0, 1.0, float(membrane), 0.0, 0.0), 1, device=dev,
                                  dtype=dtype)
    return rhs * scale + reg


def synthetic_problem(nx: int, ny: int, *, seed: int = 0, cosines: int = 64,
                      amplitude: float = 1.5, dtype: torch.dtype = torch.float64,
                      device="cuda"):
    """Feature map (a band-limited sum of ``cosines`` random 2-D cosines), a target
    (the same sum at smoothly displaced positions) and a smooth velocity field of
    about ``amplitude`` pixels: (velocity [2, ny, nx], feature, target)."""

    rng = np.random.default_rng(seed)
    x, y = np.meshgrid(np.arange(nx, dtype=np.float64), np.arange(ny, dtype=np.float64))

    def draw(n, scale):
        kx = rng.integers(-4, 5, size=n)
        ky = rng.integers(0, 5, size=n)
        amp = rng.normal(0.0, 1.0, size=n) / np.sqrt(1.0 + kx * kx + ky * ky) * scale
        ph = rng.uniform(0.0, 2.0 * np.pi, size=n)
        return amp, kx, ky, ph

    def evaluate(coeffs, px, py):
        out = np.zeros_like(px)
        for a, i, j, p in zip(*coeffs):
            out += a * np.cos(2.0 * np.pi * (i * px / nx + j * py / (2.0 * ny)) + p)
        return out

    cf = draw(cosines, 1.0)
    feature = evaluate(cf, x, y)
    ux = evaluate(draw(6, 2.0), x, y)
    uy = evaluate(draw(6, 2.0), x, y)
    target = evaluate(cf, x + ux, y + uy)
    vx = evaluate(draw(8, amplitude), x, y)
    vy = evaluate(draw(8, amplitude), x, y)
    to = dict(dtype=dtype, device=device)
    return (torch.as_tensor(np.stack([vx, vy]), **to), torch.as_tensor(feature, **to),
            torch.as_tensor(target, **to))


def recompute_vs_store(nx: int, ny: int, *, squarings: int = 6, membrane: float = 0.01,
                       dtype: torch.dtype = torch.float32, repeats: int = 20,
                       device="cuda") -> dict:
    """Forward + backward time (median ms) and workspace bytes of both schedules."""

    v, f, t = synthetic_problem(nx, ny, dtype=dtype, device=device)
    code = _lib.dtype_code(dtype)
    out = {"nx": nx, "ny": ny, "squarings": squarings, "dtype": str(dtype).replace("torch.", "")}
    for name, recompute in (("recompute", True), ("store", False)):
        times = []
        for _ in range(repeats + 2):
            w = v.clone().requires_grad_(True)
            torch.cuda.synchronize(device)
            t0 = time.perf_counter()
            e = exp_warp_ssd(w, f, t, squarings=squarings, membrane=membrane,
                             recompute=recompute)
            e.backward()
            torch.cuda.synchronize(device)
            times.append((time.perf_counter() - t0) * 1e3)
        out[name] = {
            "ms": round(float(np.median(times[2:])), 4),
            "workspace_bytes": int(_lib.lib().csg_exp_warp_ssd_workspace_bytes(
                nx, ny, squarings, code, int(not recompute))),
            "map_bytes": int((2 if recompute else squarings + 1) * 2 * nx * ny
                             * (4 if dtype == torch.float32 else 8)),
        }
    return out


def relative_l2(a: torch.Tensor, b: torch.Tensor) -> float:
    """|a - b| / |b| in float64."""

    a = a.to(torch.float64)
    b = b.to(torch.float64)
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def cosine(a: torch.Tensor, b: torch.Tensor) -> float:
    a = a.to(torch.float64).reshape(-1)
    b = b.to(torch.float64).reshape(-1)
    return float(torch.dot(a, b) / (torch.linalg.vector_norm(a) * torch.linalg.vector_norm(b)))


__all__ = ["ExpWarpSSD", "exp_warp_ssd", "exp_warp_ssd_parts", "dartel_rhs", "synthetic_problem",
           "recompute_vs_store", "relative_l2", "cosine"]
