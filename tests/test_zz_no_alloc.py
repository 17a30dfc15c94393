"""The no-allocation rotation-cost entry point against the workspace one.

csg_rotation_costs_noalloc keeps its partial sums in a fixed device array
instead of a caller workspace; its results must equal csg_rotation_costs' bit
for bit, on consecutive calls (the second smaller than the first: the array
keeps the last call's contents) and on the reference's rotation fixture.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _index(verts, faces, grid, cuda):
    import cat_surface_gpu as csg

    return csg.RotationGridIndex.from_geometry(verts, faces, grid_size=grid, margin=1,
                                               device=cuda).to(cuda)


def _costs(idx, pts, sv, tv, angles, noalloc: bool):
    """Costs, located faces and weights, and the uncovered count, through the C ABI."""

    from cat_surface_gpu import _lib

    dev = idx.device
    src = torch.as_tensor(pts, dtype=torch.float32, device=dev).contiguous()
    svd = torch.as_tensor(sv, dtype=torch.float64, device=dev).contiguous()
    tvd = torch.as_tensor(tv, dtype=torch.float64, device=dev).contiguous()
    ang = torch.as_tensor(angles, dtype=torch.float64, device=dev).reshape(-1, 3).contiguous()
    na, ns = int(ang.shape[0]), int(src.shape[0])
    costs = torch.empty(na, dtype=torch.float64, device=dev)
    loc_f = torch.empty((na, ns), dtype=torch.int32, device=dev)
    loc_b = torch.empty((na, ns, 3), dtype=torch.float32, device=dev)
    unc = torch.full((1,), 7, dtype=torch.int32, device=dev)  # must be cleared by the call
    lib = _lib.lib()
    common = (_lib.ptr(src), _lib.ptr(svd), ns, _lib.ptr(tvd), _lib.ptr(idx.target_faces),
              _lib.ptr(idx.target_triangles), _lib.ptr(idx.face_spheres()),
              _lib.ptr(idx.candidate_offsets), _lib.ptr(idx.candidate_faces),
              int(idx.grid_size), _lib.ptr(ang), na, _lib.ptr(costs), _lib.ptr(loc_f),
              _lib.ptr(loc_b))
    stream = _lib.stream_handle(dev)
    if noalloc:
        _lib.check(lib.csg_rotation_costs_noalloc(*common, _lib.ptr(unc), stream),
                   "csg_rotation_costs_noalloc")
    else:
        ws = _lib.workspace(lib.csg_rotation_costs_workspace_bytes(ns, na), dev)
        _lib.check(lib.csg_rotation_costs(*common, _lib.ptr(ws), ws.numel(), None,
                                          _lib.ptr(unc), stream), "csg_rotation_costs")
    torch.cuda.synchronize(dev)
    return costs.cpu().numpy(), loc_f.cpu().numpy(), loc_b.cpu().numpy(), int(unc.item())


def _smooth_field(p):
    return np.sin(2.0 * p[:, 0]) + np.cos(3.0 * p[:, 1]) * p[:, 2]


def test_noalloc_rotation_costs_equal_workspace_entry(cuda, golden):
    from cat_surface_gpu.stencil_builder import icosphere

    from oracle import oracle as O

    # first call: 40,962 source points against a 10,242-vertex target, 46 seeds
    tv, tf = icosphere(5)
    sv, _ = icosphere(6)
    big = _index(tv, tf, 64, cuda)
    seeds = O.official_seed_grid_angles()
    args = (sv.astype(np.float32), _smooth_field(sv), _smooth_field(tv), seeds)
    ref_big = _costs(big, *args, noalloc=False)
    got_big = _costs(big, *args, noalloc=True)
    # second call, smaller: the reference fixture (162 points, 5 angles)
    g = golden("rotation")
    small = _index(g["verts"], g["faces"], 32, cuda)
    args2 = (g["src_pts"], g["src_vals"], g["tgt_vals"], g["seeds"][:5])
    ref_small = _costs(small, *args2, noalloc=False)
    got_small = _costs(small, *args2, noalloc=True)
    for ref, got in ((ref_big, got_big), (ref_small, got_small)):
        for a, b in zip(ref[:3], got[:3]):
            np.testing.assert_array_equal(b, a)
        assert got[3] == ref[3] == 0
    np.testing.assert_array_equal(got_small[1], g["loc_faces"])
    np.testing.assert_allclose(got_small[0], g["seed_costs"][:5], rtol=1e-6)


def test_noalloc_rejects_requests_beyond_its_scratch(cuda):
    from cat_surface_gpu import _lib

    lib = _lib.lib()
    cap = int(lib.csg_rotation_costs_noalloc_max_bytes())
    assert cap == 4 << 20
    assert lib.csg_rotation_costs_workspace_bytes(163842, 204) <= cap
    assert lib.csg_rotation_costs_workspace_bytes(163842, 205) > cap
    # refused on the host before any launch (no pointers are read)
    unc = torch.zeros(1, dtype=torch.int32, device=cuda)
    null = ctypes.c_void_p(0)
    rc = lib.csg_rotation_costs_noalloc(null, null, 163842, null, null, null, null, null, null,
                                        128, null, 205, null, null, null, _lib.ptr(unc),
                                        _lib.stream_handle(cuda))
    assert rc == -1
    assert b"scratch" in lib.csg_last_error()
