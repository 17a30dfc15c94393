"""Loader and call helpers for libcsg_hip.so, the gfx950 C-ABI library.

Every device computation of this package goes through this library.  There
is no CPU or torch fallback: if the library or a HIP device is missing the
call raises.  The library is built in-tree by ``make -C csrc`` (see
``__graft_entry__.build``) and lives next to this file.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path
from typing import Any

import torch

_LIB_NAME = "libcsg_hip.so"
_HERE = Path(__file__).resolve().parent
_lib: ctypes.CDLL | None = None

_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_sz = ctypes.c_size_t
_dp = ctypes.POINTER(ctypes.c_double)

# argument signatures of the exported entry points (include/csg.h)
_SIGNATURES: dict[str, tuple[Any, list[Any]]] = {
    "csg_last_error": (ctypes.c_char_p, []),
    "csg_version": (ctypes.c_char_p, []),
    "csg_device_count": (_i, [ctypes.POINTER(_i)]),
    "csg_set_fmg_schedule": (_i, [_i]),
    "csg_set_sor_placement": (_i, [_i]),
    "csg_set_sor_force_xcd": (_i, [_i]),
    "csg_reset_sor_xcd_check": (_i, []),
    "csg_set_relax_tile": (_i, [_i, _i, _i, _i, _i]),
    "csg_fmg_tiled_levels": (_i, [_i, _i, _i, _i]),
    "csg_get_relax_tile": (_i, [_i, _i, _i, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "csg_kernel_timing": (_i, [_i]),
    "csg_kernel_timing_read": (_i, [_i, _dp, _dp, ctypes.POINTER(ctypes.c_longlong)]),
    "csg_kernel_timing_name": (ctypes.c_char_p, [_i]),
    "csg_sample_field": (_i, [_vp, _i, _i, _vp, _vp, _sz, _i, _vp, _vp]),
    "csg_composition": (_i, [_vp, _vp, _i, _i, _i, _vp, _vp]),
    "csg_composition_jacobian": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _vp, _vp, _vp]),
    "csg_jacobian_of_displacement": (_i, [_vp, _d, _i, _i, _i, _vp, _vp]),
    "csg_smalldef_jac": (_i, [_vp, _d, _i, _i, _i, _vp, _vp, _vp]),
    "csg_jac_div_smalldef": (_i, [_vp, _vp, _d, _i, _i, _i, _vp, _vp]),
    "csg_expdef_workspace_bytes": (_sz, [_i, _i, _i]),
    "csg_expdef": (_i, [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _sz, _vp]),
    "csg_objfun_workspace_bytes": (_sz, [_i, _i]),
    "csg_initialise_objfun": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _i, _vp, _vp, _vp, _vp, _sz,
                                   _vp]),
    "csg_squaring_workspace_bytes": (_sz, [_i, _i, _i]),
    "csg_squaring_update": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp,
                                 _vp, _sz, _vp]),
    "csg_regularization_operator": (_i, [_vp, _dp, _i, _i, _i, _i, _vp, _vp]),
    "csg_apply_membrane_system": (_i, [_vp, _vp, _d, _d, _d, _i, _i, _i, _vp, _vp]),
    "csg_relax_membrane": (_i, [_vp, _vp, _vp, _d, _d, _d, _i, _i, _i, _i, _vp]),
    "csg_resize_field": (_i, [_vp, _i, _i, _i, _i, _i, _i, _vp, _vp]),
    "csg_fmg2_workspace_bytes": (_sz, [_i, _i, _i, _i]),
    "csg_fmg2_membrane": (_i, [_vp, _vp, _dp, _i, _i, _vp, _i, _i, _i, _i, _i, _vp, _vp, _sz,
                               _vp]),
    "csg_dartel_step_workspace_bytes": (_sz, [_i, _i, _i, _i]),
    "csg_dartel_step": (_i, [_vp, _vp, _vp, _vp, _i, _dp, _d, _i, _i, _i, _i, _i, _i, _i, _i,
                             _vp, _vp, _vp, _sz, _vp]),
    "csg_solve_workspace_bytes": (_sz, [_i, _i, _i, _i]),
    "csg_solve_dartel_maps": (_i, [_vp, _vp, _i, _i, _i, _i, _i, _d, _i, _i, _i, _i, _i, _i,
                                   _vp, _vp, _vp, _vp, _sz, _vp]),
    "csg_exp_warp_ssd_workspace_bytes": (_sz, [_i, _i, _i, _i, _i]),
    "csg_exp_warp_ssd_forward": (_i, [_vp, _vp, _vp, _i, _i, _i, _d, _i, _i, _vp, _vp, _vp, _sz,
                                      _vp]),
    "csg_exp_warp_ssd_backward": (_i, [_vp, _vp, _vp, _i, _i, _i, _d, _i, _i, _vp, _vp, _vp,
                                       _sz, _vp]),
    # surface side
    "csg_resample_vertices": (_i, [_vp, _i, _vp, _vp, _i, _vp, _vp]),
    "csg_edge_mean_workspace_bytes": (_sz, [_i]),
    "csg_edge_sum": (_i, [_vp, _vp, _i, _vp, ctypes.POINTER(_i), _vp]),
    "csg_smooth_geometry": (_i, [_vp, _vp, _vp, _i, _i, _vp, _vp, _sz, _vp]),
    "csg_smooth_geometry_batch_workspace_bytes": (_sz, [_i, _i]),
    "csg_smooth_geometry_batch": (_i, [_vp, _i, _vp, _vp, _i, ctypes.POINTER(_i), _vp, _vp, _sz,
                                       _vp]),
    "csg_smooth_geometry_patched": (_i, [_vp, _i, _i, ctypes.POINTER(_i), _i, _i, _vp, _vp, _vp,
                                         _vp, _vp, _vp, _vp, _sz, _vp, _vp, _sz, _vp]),
    "csg_vertex_normals": (_i, [_vp, _vp, _vp, _vp, _i, _vp, _vp]),
    "csg_smooth_values_workspace_bytes": (_sz, [_i, _i, _i]),
    "csg_smooth_values": (_i, [_vp, _i, _i, _vp, _i, ctypes.POINTER(_i), _vp, _vp, _i, _dp,
                               ctypes.POINTER(_i), _i, _vp, _vp, _sz, _vp]),
    "csg_smooth_values_patched": (_i, [_vp, _i, _i, _vp, _i, ctypes.POINTER(_i), _vp, _vp, _i,
                                       _dp, ctypes.POINTER(_i), _i, _i, _i, _vp, _vp, _vp, _vp,
                                       _vp, _vp, _vp, _sz, _vp, _vp, _sz, _vp]),
    "csg_minmax_workspace_bytes": (_sz, [_i]),
    "csg_minmax_normalise": (_i, [_vp, _sz, _i, _d, _i, _vp, _vp, _sz, _vp]),
    "csg_sheet_gather": (_i, [_vp, _i, _i, _vp, _vp, _i, _vp, _vp]),
    "csg_type5_workspace_bytes": (_sz, [_i, _i, _i, _i]),
    "csg_type5_sheets": (_i, [_vp, _i, _vp, _vp, _vp, _i, _vp, _vp, _i, _vp, _i, _dp,
                              ctypes.POINTER(_i), _vp, _vp, _i, _i, _vp, _vp, _sz, _vp]),
    "csg_type5_sheets_batch_workspace_bytes": (_sz, [_i, _i, _i, _i, _i]),
    "csg_type5_sheets_batch": (_i, [_vp, _vp, _i, _i, _vp, _vp, _vp, _i, _vp, _vp, _i, _dp,
                                    ctypes.POINTER(_i), ctypes.POINTER(_vp),
                                    ctypes.POINTER(_vp), _i, _i, _vp, _vp, _sz, _vp]),
    "csg_depth_workspace_bytes": (_sz, [_i, _i, _i, _i, _i]),
    "csg_depth_potential": (_i, [_vp, _i, _i, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _i,
                                 ctypes.POINTER(_i), _vp, _i, _vp, _vp, _i, _d, _i, _d, _vp, _vp,
                                 _vp, _sz, _vp]),
    "csg_depth_potential_patched": (_i, [_vp, _i, _i, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp,
                                         _i, ctypes.POINTER(_i), _vp, _i, _vp, _vp, _i, _d, _i,
                                         _d, _vp, _vp, _vp, _vp, _sz, _vp]),
    "csg_depth_potential_patched_deferred": (_i, [_vp, _i, _i, _vp, _i, _vp, _vp, _vp, _vp, _vp,
                                                  _vp, _i, ctypes.POINTER(_i), _vp, _i, _vp, _vp,
                                                  _i, _d, _i, _d, _vp, _vp, _vp, _vp, _vp, _sz,
                                                  _vp]),
    "csg_prepare_sphere_dev": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp]),
    "csg_rotation_matrices": (_i, [_vp, _i, _vp, _vp]),
    "csg_apply_flow": (_i, [_vp, _i, _vp, _i, _i, _i, _i, _vp, _vp]),
    "csg_average_xz": (_i, [_vp, _vp, _i, _i, _dp, _vp, _vp]),
    "csg_normalise_points": (_i, [_vp, _i, _vp, _vp]),
    "csg_sphere_centroids": (_i, [ctypes.POINTER(_vp), ctypes.POINTER(_i), _i, _vp, _vp]),
    "csg_prepare_sphere": (_i, [_vp, _i, _vp, _dp, _vp, _vp, _vp]),
    # point location
    "csg_grid_index_plan_workspace_bytes": (_sz, [_i]),
    "csg_top_k": (_i, [_vp, _i, _i, _vp, _vp]),
    "csg_grid_index_build_workspace_bytes": (_sz, [_i, _i, _i]),
    "csg_grid_index_build": (_i, [_vp, _i, _vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _sz, _vp]),
    "csg_grid_index_plan": (_i, [_vp, _i, _vp, _i, _i, _i, _vp, _sz, ctypes.POINTER(_i), _vp]),
    "csg_grid_index_fill_workspace_bytes": (_sz, [_i]),
    "csg_grid_index_fill": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _sz, _vp]),
    "csg_rotation_costs_workspace_bytes": (_sz, [_i, _i]),
    "csg_rotation_cost_matrices_workspace_bytes": (_sz, [_i, _i, _i]),
    "csg_face_spheres": (_i, [_vp, _i, _vp, _vp]),
    "csg_rotation_cost_matrices": (_i, [_vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp, _i,
                                        _i, _vp, _vp, _sz, _vp, _vp]),
    "csg_rotation_map_costs_workspace_bytes": (_sz, [_i, _i, _i, _i]),
    "csg_rotation_map_costs": (_i, [_vp, _vp, _i, _i, _i, _vp, _i, _vp, _vp, _sz, _vp]),
    "csg_rotation_refine_workspace_bytes": (_sz, [_i, _i]),
    "csg_rotation_refine": (_i, [_vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp, _i, _i, _i,
                                 _d, _d, _vp, _vp, _vp, _vp, _sz, _vp, _vp]),
    "csg_rotation_costs": (_i, [_vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp, _i, _vp,
                                _vp, _vp, _vp, _sz, ctypes.POINTER(_i), _vp, _vp]),
    "csg_flag_above": (_i, [_vp, _i, _i, _vp, _vp]),
    "csg_heat_parameters": (_i, [_vp, _vp, _i, _dp, _i, _vp, _vp, _i, _vp, _vp, _vp, _vp, _sz,
                                 _vp]),
    "csg_heat_parameters_from_edge_mean": (_i, [_vp, _dp, _i, _vp, _vp, _i, _vp, _vp, _vp]),
    "csg_smooth_geometry_batch_dev": (_i, [_vp, _i, _vp, _vp, _i, _vp, _i, _vp, _vp, _sz, _vp]),
    "csg_smooth_geometry_patched_dev": (_i, [_vp, _i, _i, _vp, _i, _i, _i, _vp, _vp, _vp, _vp,
                                             _vp, _vp, _vp, _sz, _vp, _vp, _sz, _vp]),
    "csg_smooth_values_dev": (_i, [_vp, _i, _i, _vp, _i, _vp, _vp, _vp, _i, _vp, _vp, _i, _i,
                                   _vp, _vp, _sz, _vp]),
    "csg_smooth_values_patched_dev": (_i, [_vp, _i, _i, _vp, _i, _vp, _vp, _vp, _i, _vp, _vp, _i,
                                           _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _sz,
                                           _vp, _vp, _sz, _vp]),
    "csg_type5_sheets_batch_dev": (_i, [_vp, _vp, _i, _i, _vp, _vp, _vp, _i, _vp, _vp, _i, _vp,
                                        _vp, _i, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _i,
                                        _i, _vp, _vp, _sz, _vp]),
    "csg_curvature_type2": (_i, [_vp, _vp, _i, _vp, _vp, _vp, _vp]),
    "csg_curvature_sheet_workspace_bytes": (_sz, [_i, _i, _i]),
    "csg_curvature_sheet": (_i, [_vp, _i, _vp, _vp, _vp, _i, _vp, _vp, _i, _i, _d, _i, _vp, _vp,
                                 _i, _i, _vp, _vp, _sz, _vp]),
    "csg_depth_pcg_workspace_bytes": (_sz, [_i, _i, _i, _i]),
    "csg_depth_potential_pcg": (_i, [_vp, _i, _i, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _i, _d,
                                     _i, _d, _vp, _vp, ctypes.POINTER(_i), _vp, _sz, _vp]),
    "csg_rotation_costs_noalloc_max_bytes": (_sz, []),
    "csg_rotation_costs_noalloc": (_i, [_vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp, _i,
                                        _vp, _vp, _vp, _vp, _vp]),
    "csg_closest_triangles": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _i, _d, _vp, _vp, _vp,
                                   ctypes.POINTER(_i), _vp, _sz, _vp]),
    "csg_closest_triangles_workspace_bytes": (_sz, [_i]),
    "csg_stencil_sphere": (_i, [_vp, _i, _i, _vp, _i, _vp, _vp, _vp, _vp]),
    "csg_resample_unit": (_i, [_vp, _vp, _vp, _i, _vp, _vp, _vp]),
    "csg_unit_points64": (_i, [_vp, _i, _vp, _vp]),
    "csg_closest_triangles_pruned": (_i, [_vp, _i, _vp, _vp, _i, _vp, _vp, _i, _d, _vp, _vp, _vp,
                                          _vp, ctypes.POINTER(_i), _vp, _sz, _vp]),
    "csg_closest_triangles_block": (_i, [_vp, _i, _vp, _vp, _i, _vp, _vp, _i, _i, _vp, _vp, _vp,
                                         _vp, ctypes.POINTER(_i), _vp, _sz, _vp]),
}


class HipLibraryError(RuntimeError):
    """The HIP library is missing or a call into it failed at run time."""


def library_path() -> Path:
    override = os.environ.get("CSG_HIP_LIBRARY")
    return Path(override) if override else _HERE / _LIB_NAME


def lib() -> ctypes.CDLL:
    """Load libcsg_hip.so once; raise if it is not built."""

    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not path.is_file():
        raise HipLibraryError(
            f"{_LIB_NAME} is not built at {path}; run `make -C cat-surface-gpu_amd/csrc` "
            "(or __graft_entry__.build()).  There is no CPU fallback."
        )
    handle = ctypes.CDLL(str(path))
    for name, (restype, argtypes) in _SIGNATURES.items():
        fn = getattr(handle, name, None)
        if fn is None:
            continue
        fn.restype = restype
        fn.argtypes = argtypes
    _lib = handle
    return handle


def exported_symbols() -> list[str]:
    return list(_SIGNATURES)


def check(code: int, what: str) -> None:
    """Map a C-ABI status code to the reference's exception types."""

    if code == 0:
        return
    message = lib().csg_last_error().decode("utf-8", "replace")
    text = f"{what}: {message}"
    if code == -1:
        raise ValueError(text)
    if code == -3:
        raise NotImplementedError(text)
    raise HipLibraryError(text)


def dtype_code(dtype: torch.dtype) -> int:
    if dtype == torch.float32:
        return 0
    if dtype == torch.float64:
        return 1
    raise ValueError(f"only float32 and float64 are supported, got {dtype}")


def stream_handle(device: torch.device) -> int:
    return int(torch.cuda.current_stream(device).cuda_stream)


def ptr(tensor: torch.Tensor | None) -> int | None:
    if tensor is None:
        return None
    return int(tensor.data_ptr())


def workspace(nbytes: int, device: torch.device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)


def doubles(values) -> Any:
    arr = (ctypes.c_double * len(values))(*[float(v) for v in values])
    return arr
