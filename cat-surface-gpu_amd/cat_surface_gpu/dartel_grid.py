"""CAT-Surface's two-dimensional DARTEL operators on MI355X (HIP, gfx950).

Drop-in for ``cat_surface_gpu.dartel_grid`` of the reference
(src/cat_surface_gpu/dartel_grid.py): same function names, argument meaning,
shapes ``[C, ny, nx]`` (one-based maps) and error types.  Every operator runs
in a hand-written HIP kernel of libcsg_hip.so; the ``kernel=`` seam accepts
``"auto"`` and ``"hip"``.  ``"torch"`` and ``"triton"`` name backends that are
not part of this build and raise ``ValueError`` (the reference's explicit
no-fallback contract, dartel_grid.py:259-266).  There is no CPU path.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Sequence

import torch

from . import _lib

_KERNELS = {"auto", "hip"}


class DeviceUnavailable(RuntimeError):
    """Report that an explicitly requested Torch device is unavailable."""


@dataclass(frozen=True)
class GridSpec:
    """Describe the x and y dimensions of a DARTEL grid."""

    nx: int
    ny: int

    def __post_init__(self) -> None:
        if self.nx < 1 or self.ny < 1:
            raise ValueError(f"Grid dimensions must be positive, got {(self.nx, self.ny)}")

    @property
    def points(self) -> int:
        return self.nx * self.ny

    @classmethod
    def from_shape(cls, shape: Sequence[int]) -> "GridSpec":
        if len(shape) != 3 or shape[0] != 2:
            raise ValueError(f"Grid shape must be [2, ny, nx], got {tuple(shape)}")
        return cls(nx=int(shape[2]), ny=int(shape[1]))


def resolve_device(device: str | torch.device = "auto") -> torch.device:
    """Resolve the HIP device; this build has no CPU path.

    Mirrors dartel_grid.py:67-87: an unavailable explicit device raises
    ``DeviceUnavailable`` instead of falling back.
    """

    if isinstance(device, str) and device == "auto":
        if not torch.cuda.is_available():
            raise DeviceUnavailable(
                "no HIP device is visible; this MI355X build has no CPU fallback"
            )
        return torch.device("cuda", torch.cuda.current_device())
    resolved = torch.device(device)
    if resolved.type != "cuda":
        raise DeviceUnavailable(
            f"this MI355X build runs only on HIP devices ('cuda' in torch), got {resolved}"
        )
    if not torch.cuda.is_available():
        raise DeviceUnavailable("HIP device requested but no GPU is visible")
    if resolved.index is not None and resolved.index >= torch.cuda.device_count():
        raise DeviceUnavailable(
            f"HIP device {resolved.index} requested, {torch.cuda.device_count()} visible"
        )
    if resolved.index is None:
        resolved = torch.device("cuda", torch.cuda.current_device())
    return resolved


def _check_kernel(kernel: str) -> None:
    if kernel in _KERNELS:
        return
    if kernel in {"torch", "triton"}:
        raise ValueError(
            f"kernel={kernel!r} is not part of this build; use kernel='hip' (or 'auto')"
        )
    raise ValueError(f"kernel must be 'auto' or 'hip', got {kernel!r}")


def _dev(value, device: torch.device, dtype: torch.dtype) -> torch.Tensor:
    return torch.as_tensor(value, dtype=dtype, device=device).contiguous()


def _as_grid(value, spec: GridSpec, device, dtype, name: str) -> torch.Tensor:
    tensor = _dev(value, device, dtype)
    expected = (2, spec.ny, spec.nx)
    if tuple(tensor.shape) != expected:
        raise ValueError(f"{name} shape must be {expected}, got {tuple(tensor.shape)}")
    return tensor


def _as_jacobian(value, spec: GridSpec, device, dtype, name: str) -> torch.Tensor:
    tensor = _dev(value, device, dtype)
    expected = (4, spec.ny, spec.nx)
    if tuple(tensor.shape) != expected:
        raise ValueError(f"{name} shape must be {expected}, got {tuple(tensor.shape)}")
    return tensor


def _call(name: str, *args) -> None:
    _lib.check(getattr(_lib.lib(), name)(*args), name)


def from_c_layout(value, nx: int, ny: int, components: int = 2, *,
                  device: str | torch.device = "auto", dtype: torch.dtype = torch.float64):
    spec = GridSpec(nx=nx, ny=ny)
    target = resolve_device(device)
    tensor = torch.as_tensor(value, dtype=dtype, device=target)
    expected = components * spec.points
    if tensor.numel() != expected:
        raise ValueError(f"CAT C layout requires {expected} elements, got {tensor.numel()}")
    return tensor.reshape(components, spec.ny, spec.nx).contiguous()


def to_c_layout(value: torch.Tensor) -> torch.Tensor:
    if value.ndim != 3:
        raise ValueError(f"Input must be a three-dimensional tensor, got {tuple(value.shape)}")
    return value.contiguous().reshape(-1)


def make_identity_map(spec: GridSpec, *, device: str | torch.device = "auto",
                      dtype: torch.dtype = torch.float64) -> torch.Tensor:
    """CAT-Surface's one-based identity coordinate map (dartel_grid.py:174-187)."""

    target = resolve_device(device)
    zero = torch.zeros((2, spec.ny, spec.nx), dtype=dtype, device=target)
    out = torch.empty_like(zero)
    _call("csg_smalldef_jac", _lib.ptr(zero), 0.0, spec.nx, spec.ny, _lib.dtype_code(dtype),
          _lib.ptr(out), _lib.ptr(torch.empty((4, spec.ny, spec.nx), dtype=dtype, device=target)),
          _lib.stream_handle(target))
    return out


def sample_field(field, x, y, *, device: str | torch.device = "auto",
                 dtype: torch.dtype = torch.float64) -> torch.Tensor:
    """Bilinear sample of a 2-D field (dartel_grid.py:510-531)."""

    target = resolve_device(device)
    f = _dev(field, target, dtype)
    if f.ndim != 2:
        raise ValueError(f"sample_field accepts a 2D scalar field, got {tuple(f.shape)}")
    xt = _dev(x, target, dtype)
    yt = _dev(y, target, dtype)
    if xt.shape != yt.shape:
        raise ValueError("sample_field x and y coordinates must have the same shape")
    out = torch.empty_like(xt)
    _call("csg_sample_field", _lib.ptr(f), int(f.shape[1]), int(f.shape[0]), _lib.ptr(xt),
          _lib.ptr(yt), xt.numel(), _lib.dtype_code(dtype), _lib.ptr(out),
          _lib.stream_handle(target))
    return out


def composition(a, b, *, device: str | torch.device = "auto", dtype: torch.dtype = torch.float64,
                optimized: bool = True) -> torch.Tensor:
    """Compose two coordinate maps, b∘a (dartel_grid.py:552-575)."""

    target = resolve_device(device)
    spec = GridSpec.from_shape(torch.as_tensor(a).shape)
    at = _as_grid(a, spec, target, dtype, "A")
    bt = _as_grid(b, spec, target, dtype, "B")
    out = torch.empty_like(at)
    _call("csg_composition", _lib.ptr(at), _lib.ptr(bt), spec.nx, spec.ny,
          _lib.dtype_code(dtype), _lib.ptr(out), _lib.stream_handle(target))
    return out


def composition_jacobian(a, ja, b, jb, *, device: str | torch.device = "auto",
                         dtype: torch.dtype = torch.float64, optimized: bool = True):
    """(b∘a, (Jb∘a)·Ja) (dartel_grid.py:578-621)."""

    target = resolve_device(device)
    spec = GridSpec.from_shape(torch.as_tensor(a).shape)
    at = _as_grid(a, spec, target, dtype, "A")
    bt = _as_grid(b, spec, target, dtype, "B")
    jat = _as_jacobian(ja, spec, target, dtype, "JA")
    jbt = _as_jacobian(jb, spec, target, dtype, "JB")
    out_map = torch.empty_like(at)
    out_jac = torch.empty_like(jat)
    _call("csg_composition_jacobian", _lib.ptr(at), _lib.ptr(jat), _lib.ptr(bt), _lib.ptr(jbt),
          spec.nx, spec.ny, _lib.dtype_code(dtype), _lib.ptr(out_map), _lib.ptr(out_jac),
          _lib.stream_handle(target))
    return out_map, out_jac


def jacobian_of_displacement(displacement, scale: float = 1.0, *,
                             device: str | torch.device = "auto",
                             dtype: torch.dtype = torch.float64) -> torch.Tensor:
    """I + scale·∇v, channels [∂xφx, ∂xφy, ∂yφx, ∂yφy] (dartel_grid.py:624-652)."""

    target = resolve_device(device)
    spec = GridSpec.from_shape(torch.as_tensor(displacement).shape)
    v = _as_grid(displacement, spec, target, dtype, "displacement")
    out = torch.empty((4, spec.ny, spec.nx), dtype=dtype, device=target)
    _call("csg_jacobian_of_displacement", _lib.ptr(v), float(scale), spec.nx, spec.ny,
          _lib.dtype_code(dtype), _lib.ptr(out), _lib.stream_handle(target))
    return out


def smalldef_jac(displacement, scale: float, *, device: str | torch.device = "auto",
                 dtype: torch.dtype = torch.float64):
    """(id + scale·v, I + scale·∇v) (dartel_grid.py:655-676)."""

    target = resolve_device(device)
    spec = GridSpec.from_shape(torch.as_tensor(displacement).shape)
    v = _as_grid(displacement, spec, target, dtype, "displacement")
    out_map = torch.empty_like(v)
    out_jac = torch.empty((4, spec.ny, spec.nx), dtype=dtype, device=target)
    _call("csg_smalldef_jac", _lib.ptr(v), float(scale), spec.nx, spec.ny,
          _lib.dtype_code(dtype), _lib.ptr(out_map), _lib.ptr(out_jac), _lib.stream_handle(target))
    return out_map, out_jac


def jac_div_smalldef(jacobian, displacement, scale: float, *,
                     device: str | torch.device = "auto",
                     dtype: torch.dtype = torch.float64) -> torch.Tensor:
    """J·(I + scale·∇v)^-1 (dartel_grid.py:679-720)."""

    target = resolve_device(device)
    spec = GridSpec.from_shape(torch.as_tensor(displacement).shape)
    v = _as_grid(displacement, spec, target, dtype, "displacement")
    jac = _as_jacobian(jacobian, spec, target, dtype, "jacobian")
    out = torch.empty_like(jac)
    _call("csg_jac_div_smalldef", _lib.ptr(jac), _lib.ptr(v), float(scale), spec.nx, spec.ny,
          _lib.dtype_code(dtype), _lib.ptr(out), _lib.stream_handle(target))
    return out


def initialise_objfun(source, target, transformation, jacobian, distortion=None, *,
                      device: str | torch.device = "auto", dtype: torch.dtype = torch.float64):
    """Objective ½Σd²w, gradient and Gauss-Newton Hessian (dartel_grid.py:723-775).

    The objective is accumulated in float64 for either dtype and returned as a
    0-d tensor of ``dtype``.
    """

    dev = resolve_device(device)
    src = _dev(source, dev, dtype)
    tgt = _dev(target, dev, dtype)
    if src.ndim != 2 or tgt.ndim != 2:
        raise ValueError("initialise_objfun source and target must be 2D scalar fields")
    if tuple(src.shape) != tuple(tgt.shape):
        raise ValueError("initialise_objfun source and target must have the same shape")
    spec = GridSpec(nx=int(src.shape[1]), ny=int(src.shape[0]))
    phi = _as_grid(transformation, spec, dev, dtype, "transformation")
    jac = _as_jacobian(jacobian, spec, dev, dtype, "jacobian")
    dist = None
    if distortion is not None:
        dist = _dev(distortion, dev, dtype)
        if tuple(dist.shape) != (spec.ny, spec.nx):
            raise ValueError(f"Distortion shape must be {(spec.ny, spec.nx)}")
    rhs = torch.empty((2, spec.ny, spec.nx), dtype=dtype, device=dev)
    hes = torch.empty((3, spec.ny, spec.nx), dtype=dtype, device=dev)
    objective = torch.empty((), dtype=torch.float64, device=dev)
    ws = _lib.workspace(_lib.lib().csg_objfun_workspace_bytes(spec.nx, spec.ny), dev)
    _call("csg_initialise_objfun", _lib.ptr(src), _lib.ptr(tgt), _lib.ptr(phi), _lib.ptr(jac),
          _lib.ptr(dist), spec.nx, spec.ny, _lib.dtype_code(dtype), _lib.ptr(rhs), _lib.ptr(hes),
          _lib.ptr(objective), _lib.ptr(ws), ws.numel(), _lib.stream_handle(dev))
    return objective.to(dtype), rhs, hes


def squaring_update(right_hand, hessian, transformation, jacobian, k: int, *,
                    save_transformation: bool = False, return_scratch: bool = False,
                    device: str | torch.device = "auto", dtype: torch.dtype = torch.float64,
                    optimized: bool = True, kernel: str = "auto"):
    """k fused squaring updates of (b, A) along the composed map (dartel_grid.py:778-914)."""

    if k < 0:
        raise ValueError(f"squaring step count k must be non-negative, got {k}")
    dev = resolve_device(device)
    _check_kernel(kernel)
    spec = GridSpec.from_shape(torch.as_tensor(transformation).shape)
    b = _as_grid(right_hand, spec, dev, dtype, "right_hand")
    a = _dev(hessian, dev, dtype)
    if tuple(a.shape) != (3, spec.ny, spec.nx):
        raise ValueError(f"Hessian shape must be {(3, spec.ny, spec.nx)}")
    phi = _as_grid(transformation, spec, dev, dtype, "transformation")
    jac = _as_jacobian(jacobian, spec, dev, dtype, "jacobian")
    b_out = torch.empty_like(b)
    a_out = torch.empty_like(a)
    m_out = torch.empty_like(phi)
    j_out = torch.empty_like(jac)
    s_out = torch.empty_like(phi)
    code = _lib.dtype_code(dtype)
    ws = _lib.workspace(_lib.lib().csg_squaring_workspace_bytes(spec.nx, spec.ny, code), dev)
    _call("csg_squaring_update", _lib.ptr(b), _lib.ptr(a), _lib.ptr(phi), _lib.ptr(jac), int(k),
          int(bool(save_transformation)), spec.nx, spec.ny, code, _lib.ptr(b_out), _lib.ptr(a_out),
          _lib.ptr(m_out), _lib.ptr(j_out), _lib.ptr(s_out), _lib.ptr(ws), ws.numel(),
          _lib.stream_handle(dev))
    result = (b_out, a_out, m_out, j_out)
    if return_scratch:
        return result + (s_out,)
    return result


def regularization_operator(field, params: Sequence[float], rtype: int = 1, *,
                            device: str | torch.device = "auto",
                            dtype: torch.dtype = torch.float64) -> torch.Tensor:
    """L v for rtype 0 (linear elastic), 1 (membrane), 2 (bending) (dartel_grid.py:917-1015)."""

    if len(params) != 5:
        raise ValueError(f"regularization parameters must contain 5 values, got {len(params)}")
    if rtype not in {0, 1, 2}:
        raise ValueError(f"regularization type must be 0, 1, or 2, got {rtype}")
    dev = resolve_device(device)
    spec = GridSpec.from_shape(torch.as_tensor(field).shape)
    v = _as_grid(field, spec, dev, dtype, "field")
    out = torch.empty_like(v)
    _call("csg_regularization_operator", _lib.ptr(v), _lib.doubles(params), int(rtype), spec.nx,
          spec.ny, _lib.dtype_code(dtype), _lib.ptr(out), _lib.stream_handle(dev))
    return out


def _membrane_weights(params: Sequence[float]) -> tuple[float, float, float]:
    sx, sy, mu, _lam, identity = (float(item) for item in params)
    return (mu * (2.0 * sx * sx + 2.0 * sy * sy) + identity, -mu * sy * sy, -mu * sx * sx)


def apply_membrane_system(hessian, field, params: Sequence[float], *, kernel: str = "auto",
                          device: str | torch.device = "auto",
                          dtype: torch.dtype = torch.float64) -> torch.Tensor:
    """(L_membrane + A)·u (dartel_grid.py:1018-1070)."""

    dev = resolve_device(device)
    f = torch.as_tensor(field)
    spec = GridSpec(nx=int(f.shape[-1]), ny=int(f.shape[-2]))
    a = _dev(hessian, dev, dtype)
    if tuple(a.shape) != (3, spec.ny, spec.nx):
        raise ValueError(f"Hessian shape must be {(3, spec.ny, spec.nx)}")
    u = _as_grid(field, spec, dev, dtype, "field")
    _check_kernel(kernel)
    w00, w01, w10 = _membrane_weights(params)
    out = torch.empty_like(u)
    _call("csg_apply_membrane_system", _lib.ptr(a), _lib.ptr(u), w00, w01, w10, spec.nx, spec.ny,
          _lib.dtype_code(dtype), _lib.ptr(out), _lib.stream_handle(dev))
    return out


def relax_membrane(hessian, right_hand, params: Sequence[float], nit: int, initial=None, *,
                   kernel: str = "auto", device: str | torch.device = "auto",
                   dtype: torch.dtype = torch.float64) -> torch.Tensor:
    """nit red-black block Gauss-Seidel iterations (dartel_grid.py:1073-1208).

    Odd widths follow the reference's sequential CPU order exactly (the
    reference's own GPU path rejects them); see k_relax_color.
    """

    if nit < 0:
        raise ValueError(f"relaxation count must be non-negative, got {nit}")
    dev = resolve_device(device)
    b = _dev(right_hand, dev, dtype)
    if b.ndim != 3 or b.shape[0] != 2:
        raise ValueError("right_hand must have shape [2, ny, nx]")
    spec = GridSpec(nx=int(b.shape[2]), ny=int(b.shape[1]))
    a = _dev(hessian, dev, dtype)
    if tuple(a.shape) != (3, spec.ny, spec.nx):
        raise ValueError(f"Hessian shape must be {(3, spec.ny, spec.nx)}")
    if initial is None:
        solution = torch.zeros_like(b)
    else:
        solution = _as_grid(initial, spec, dev, dtype, "initial").clone()
    if len(params) != 5:
        raise ValueError(f"regularization parameters must contain 5 values, got {len(params)}")
    _check_kernel(kernel)
    w00, w01, w10 = _membrane_weights(params)
    _call("csg_relax_membrane", _lib.ptr(a), _lib.ptr(b), _lib.ptr(solution), w00, w01, w10,
          int(nit), spec.nx, spec.ny, _lib.dtype_code(dtype), _lib.stream_handle(dev))
    return solution


def resize_field(field, target: GridSpec, *, kernel: str = "auto",
                 device: str | torch.device = "auto",
                 dtype: torch.dtype = torch.float64) -> torch.Tensor:
    """Quadratic B-spline resize (dartel_grid.py:239-305)."""

    dev = resolve_device(device)
    value = torch.as_tensor(field, dtype=dtype, device=dev)
    if value.ndim == 2:
        value = value.unsqueeze(0)
    if value.ndim != 3:
        raise ValueError(f"resize_field expects [channel, y, x], got {tuple(value.shape)}")
    _check_kernel(kernel)
    value = value.contiguous()
    out = torch.empty((value.shape[0], target.ny, target.nx), dtype=dtype, device=dev)
    _call("csg_resize_field", _lib.ptr(value), int(value.shape[0]), int(value.shape[2]),
          int(value.shape[1]), target.nx, target.ny, _lib.dtype_code(dtype), _lib.ptr(out),
          _lib.stream_handle(dev))
    return out


def set_fmg_schedule(fused: bool) -> bool:
    """Choose the multigrid schedule of every later FMG call; returns the previous one.

    ``True`` (the default) relaxes the fine levels on LDS tiles and the coarse
    levels in one workgroup per job; ``False`` launches one kernel per reference
    operator call (dartel_grid.py:1211-1417).  Both compute identical values.
    """

    return bool(_lib.lib().csg_set_fmg_schedule(1 if fused else 0))


def fmg2_membrane(hessian, right_hand, params: Sequence[float], cycles: int = 3, nit: int = 3,
                  initial=None, *, kernel: str = "auto", device: str | torch.device = "auto",
                  dtype: torch.dtype = torch.float64,
                  assume_initial_nonzero: bool = False) -> torch.Tensor:
    """Full-multigrid membrane solve (dartel_grid.py:1211-1417)."""

    if cycles < 0 or nit < 0:
        raise ValueError(f"FMG cycles and nit must be non-negative, got {(cycles, nit)}")
    dev = resolve_device(device)
    b = _dev(right_hand, dev, dtype)
    if b.ndim != 3 or b.shape[0] != 2:
        raise ValueError("right_hand must have shape [2, ny, nx]")
    spec = GridSpec(nx=int(b.shape[2]), ny=int(b.shape[1]))
    a = _dev(hessian, dev, dtype)
    if tuple(a.shape) != (3, spec.ny, spec.nx):
        raise ValueError(f"Hessian shape must be {(3, spec.ny, spec.nx)}")
    if len(params) != 5:
        raise ValueError(f"regularization parameters must contain 5 values, got {len(params)}")
    _check_kernel(kernel)
    init = None
    coarse_init = 1
    if initial is not None:
        init = _as_grid(initial, spec, dev, dtype, "initial")
        if assume_initial_nonzero:
            coarse_init = 0
        else:
            coarse_init = int(not bool(torch.any(init).item()))
    out = torch.empty_like(b)
    code = _lib.dtype_code(dtype)
    ws = _lib.workspace(_lib.lib().csg_fmg2_workspace_bytes(spec.nx, spec.ny, 1, code), dev)
    _call("csg_fmg2_membrane", _lib.ptr(a), _lib.ptr(b), _lib.doubles(params), int(cycles),
          int(nit), _lib.ptr(init), coarse_init, spec.nx, spec.ny, 1, code, _lib.ptr(out),
          _lib.ptr(ws), ws.numel(), _lib.stream_handle(dev))
    return out


def dartel_step(source, target, velocity, *, k: int, rtype: int = 1,
                params: Sequence[float] = (1.0, 1.0, 0.125, 0.0, 0.0), lmreg: float = 1e-3,
                cycles: int = 3, nit: int = 3, code: int = 1, distortion=None,
                kernel: str = "auto", squaring_kernel: str | None = None,
                device: str | torch.device = "auto", dtype: torch.dtype = torch.float64,
                optimized: bool = True, assume_initial_nonzero: bool = False):
    """One membrane-regularised Gauss-Newton DARTEL step (dartel_grid.py:1420-1616).

    Returns (updated velocity, metrics[3] = (objective, ½Σ Lv·v, ‖rhs‖)).
    """

    if k < 0:
        raise ValueError(f"DARTEL step count k must be non-negative, got {k}")
    if cycles < 0 or nit < 0:
        raise ValueError(f"DARTEL cycles and nit must be non-negative, got {(cycles, nit)}")
    if squaring_kernel not in {None, "auto", "hip", "torch", "triton"}:
        raise ValueError("squaring_kernel must be None, 'auto', or 'hip'")
    if code not in {0, 1}:
        raise NotImplementedError("The DARTEL implementation supports only code 0 and 1")
    if rtype != 1:
        raise NotImplementedError("The DARTEL implementation supports only rtype 1")
    if len(params) != 5:
        raise ValueError(
            f"DARTEL regularization parameters must contain 5 values, got {len(params)}"
        )
    _check_kernel(kernel)
    if squaring_kernel is not None:
        _check_kernel(squaring_kernel)
    dev = resolve_device(device)
    src = _dev(source, dev, dtype)
    tgt = _dev(target, dev, dtype)
    if src.ndim != 2 or tgt.ndim != 2:
        raise ValueError("DARTEL source and target must be 2D scalar fields")
    if tuple(src.shape) != tuple(tgt.shape):
        raise ValueError("DARTEL source and target must have the same shape")
    spec = GridSpec(nx=int(src.shape[1]), ny=int(src.shape[0]))
    v = _as_grid(velocity, spec, dev, dtype, "velocity")
    dist = None
    if distortion is not None:
        dist = _dev(distortion, dev, dtype)
        if dist.numel() != spec.points:
            raise ValueError(f"Distortion must be {spec.points}")
        dist = dist.reshape(spec.ny, spec.nx).contiguous()
    out = torch.empty_like(v)
    metrics = torch.empty(3, dtype=torch.float64, device=dev)
    dcode = _lib.dtype_code(dtype)
    ws = _lib.workspace(_lib.lib().csg_dartel_step_workspace_bytes(spec.nx, spec.ny, 1, dcode),
                        dev)
    _call("csg_dartel_step", _lib.ptr(src), _lib.ptr(tgt), _lib.ptr(v), _lib.ptr(dist), int(k),
          _lib.doubles(params), float(lmreg), int(cycles), int(nit), int(code),
          int(bool(assume_initial_nonzero)), spec.nx, spec.ny, 1, dcode, _lib.ptr(out),
          _lib.ptr(metrics), _lib.ptr(ws), ws.numel(), _lib.stream_handle(dev))
    return out, metrics.to(dtype)


def expdef(displacement, k: int, *, return_jacobian: bool = False,
           device: str | torch.device = "auto", dtype: torch.dtype = torch.float64,
           optimized: bool = True):
    """Exponentiate a velocity field by scaling and squaring (dartel_grid.py:1619-1666)."""

    if k < 0:
        raise ValueError(f"exponentiation step count k must be non-negative, got {k}")
    dev = resolve_device(device)
    spec = GridSpec.from_shape(torch.as_tensor(displacement).shape)
    v = _as_grid(displacement, spec, dev, dtype, "displacement")
    code = _lib.dtype_code(dtype)
    out_map = torch.empty_like(v)
    out_jac = (torch.empty((4, spec.ny, spec.nx), dtype=dtype, device=dev)
               if return_jacobian else None)
    ws = _lib.workspace(_lib.lib().csg_expdef_workspace_bytes(spec.nx, spec.ny, code), dev)
    _call("csg_expdef", _lib.ptr(v), spec.nx, spec.ny, code, int(k), _lib.ptr(out_map),
          _lib.ptr(out_jac), _lib.ptr(ws), ws.numel(), _lib.stream_handle(dev))
    if return_jacobian:
        return out_map, out_jac
    return out_map
