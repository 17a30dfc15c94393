"""HIP surface-side operators vs the reference golden vectors and the oracle.

Float32 geometry paths are compared within float32 rounding (the reference's
sums are torch CPU reductions whose order is not specified); float64 value
paths within 1e-12.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small(golden, cuda):
    import cat_surface_gpu as csg

    g = golden("surface")
    st = csg.SurfaceStencil(
        sphere_points=g["verts"], faces=g["faces"], surface_indices=g["surf_idx"],
        surface_weights=g["surf_w"], sheet_indices=g["sheet_idx"], sheet_weights=g["sheet_w"],
        nx=32, ny=16, source_points=g["verts"].shape[0],
        unit_sphere_points=g["verts"].astype(np.float32))
    return g, st.to(cuda)


def _np(t):
    return t.detach().cpu().numpy()


def test_resample_and_sheet_gather(small, cuda):
    g, dev = small
    surf = torch.as_tensor(g["surface"], device=cuda)
    np.testing.assert_array_equal(_np(dev.resample_vertices(surf)), g["resampled"])
    np.testing.assert_allclose(_np(dev.resample_vertices(surf.double())), g["resampled64"],
                               rtol=1e-15)
    vals = torch.as_tensor(g["values"], device=cuda)
    np.testing.assert_allclose(_np(dev.map_values_to_sheet(vals)), g["sheet_of_values"],
                               rtol=1e-15, atol=0)


def test_heat_smoothing_matches_golden(small, cuda):
    g, dev = small
    surf = torch.as_tensor(g["surface"], device=cuda)
    assert abs(dev.edge_mean_distance(surf) - float(g["edge_mean"])) < 1e-12
    for fwhm, key in ((50.0, "smooth_geom50"), (15.0, "smooth_geom15")):
        np.testing.assert_allclose(_np(dev.smooth_geometry(surf, fwhm)), g[key], rtol=2e-7,
                                   atol=1e-5)
    vals = torch.as_tensor(g["values"], device=cuda)
    np.testing.assert_allclose(_np(dev.smooth_values(vals, surf, 5.0)), g["smooth_vals5"],
                               rtol=1e-12, atol=1e-14)
    many = dev.smooth_values_many(vals, surf, (5.0, 5.0 / 3.0))
    np.testing.assert_allclose(_np(many[0]), g["smooth_many0"], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(_np(many[1]), g["smooth_many1"], rtol=1e-12, atol=1e-14)
    # problems on different meshes of one topology in one call equal separate calls
    other = surf * 1.25
    params = [(7, 1.5), (3, 0.5), (5, 2.0)]
    rows = torch.stack([vals, vals * 2.0, vals - 1.0]).double()
    batched = dev._smooth_values_batch(rows, torch.stack([surf, other]), params, False,
                                       mesh_of=(1, 0, 1))
    for b, (mesh, p) in enumerate(zip((other, surf, other), params)):
        single = dev._smooth_values_batch(rows[b:b + 1], mesh, [p], False)[0]
        assert torch.equal(batched[b], single)


def test_normals_and_type5_sheets(small, cuda):
    g, dev = small
    surf = torch.as_tensor(g["surface"], device=cuda)
    np.testing.assert_allclose(_np(dev.vertex_normals(surf)), g["normals"], atol=2e-6)
    maps = dev.curvature_type5_to_sheet_many(surf, (5.0 / 9.0, 5.0 / 27.0))
    np.testing.assert_allclose(_np(maps[0]), g["type5_sheet0"], atol=2e-6)
    np.testing.assert_allclose(_np(maps[1]), g["type5_sheet1"], atol=2e-6)
    f32 = dev.curvature_type5_to_sheet_many(surf, (5.0 / 9.0,), out_dtype=torch.float32)
    np.testing.assert_allclose(_np(f32[0]), g["type5_sheet0"], atol=3e-6)


def test_type2_curvature_and_single_fwhm_sheets(small, cuda):
    """Type-2 curvature (surface_stencil.py:552-622) and curvature_to_sheet for
    both curvtypes (:624-644).  Given the reference's float32 normals the HIP
    shape-operator fit equals the oracle's to float64 rounding (the kernel
    contracts products into FMAs: 1e-14 relative); with its own normals (2e-6 from
    torch's) the normalised values and sheets stay within 5e-6 of the reference."""

    from oracle import oracle as O

    from cat_surface_gpu import _lib

    g, dev = small
    surf = torch.as_tensor(g["surface"], device=cuda)
    n = int(surf.shape[0])
    topo = dev.topology
    normals = torch.as_tensor(g["normals"], device=cuda).contiguous()
    raw = torch.empty(n, dtype=torch.float64, device=cuda)
    _lib.check(_lib.lib().csg_curvature_type2(_lib.ptr(surf), _lib.ptr(normals), n,
                                              _lib.ptr(topo.nbr_off), _lib.ptr(topo.nbr),
                                              _lib.ptr(raw), _lib.stream_handle(cuda)),
               "csg_curvature_type2")
    import ctypes

    off, nbr = O.neighbour_csr(g["faces"], n)
    want = np.empty(n)
    O.lib().oracle_curvature_type2(O._f(np.ascontiguousarray(g["surface"])),
                                   O._f(np.ascontiguousarray(g["normals"])), O._i(off),
                                   O._i(nbr), ctypes.c_int(n), O._d(want))
    np.testing.assert_allclose(_np(raw), want, rtol=1e-13, atol=0)
    np.testing.assert_allclose(_np(dev.curvature_type2(surf, 5.0)), g["type2_values"], rtol=0,
                               atol=5e-6)
    for dt, tol in ((torch.float64, 5e-6), (torch.float32, 6e-6)):
        t2 = dev.curvature_to_sheet(surf, 5.0 / 81.0, curvtype=2, out_dtype=dt)
        assert t2.dtype == dt and tuple(t2.shape) == (16, 32)
        np.testing.assert_allclose(_np(t2), g["type2_sheet"], rtol=0, atol=tol)
        t5 = dev.curvature_to_sheet(surf, 5.0 / 9.0, curvtype=5, out_dtype=dt)
        np.testing.assert_allclose(_np(t5), g["type5_sheet_single"], rtol=0, atol=tol)
    # the per-step rule of surface_warp.py:427-448
    both5 = dev.curvature_sheets(surf, (5.0 / 9.0, 5.0 / 27.0), (5, 5))
    np.testing.assert_allclose(_np(both5[0]), g["type5_sheet0"], atol=2e-6)
    mixed = dev.curvature_sheets(surf, (5.0 / 9.0, 5.0 / 81.0), (5, 2))
    np.testing.assert_allclose(_np(mixed[0]), g["type5_sheet_single"], atol=5e-6)
    np.testing.assert_allclose(_np(mixed[1]), g["type2_sheet"], atol=5e-6)
    with pytest.raises(NotImplementedError):
        dev.curvature_to_sheet(surf, 1.0, curvtype=3)


def test_device_heat_parameters_equal_the_host_formula(small, cuda):
    """csg_heat_parameters (edge mean, iterations and sigma on the device) equals
    heat_parameters_from_edge_mean on the host bit for bit, the iteration counts
    whose host square root (libm pow) is not the correctly rounded one included."""

    import math

    from cat_surface_gpu.surface_stencil import (EdgeHeat, edge_heat, edge_means,
                                                 heat_from_edge_mean,
                                                 heat_parameters_from_edge_mean)

    g, dev = small
    surf = torch.as_tensor(g["surface"], device=cuda)
    fws = (50.0, 15.0, 5.0 / 9.0)
    h = edge_heat([dev], [surf], [fws])[0]
    em = edge_means([dev], [surf])[0]
    assert abs(em - float(g["edge_mean"])) <= 1e-12 * em
    assert float(h.edge_mean.item()) == em
    assert h.host() == [heat_parameters_from_edge_mean(em, f) for f in fws]
    # counts where k ** 0.5 != sqrt(k) (2921, 3541, 5579) and ordinary ones
    counts = [1, 2, 44, 501, 2921, 3541, 5579, 7827, 2 ** 20 - 1]
    assert any(k ** 0.5 != math.sqrt(k) for k in counts)
    e = 1.7
    fw = [math.sqrt((k - 0.5) * e * e / 0.541011) for k in counts]
    base = EdgeHeat((), h.iters[:0], h.sigma[:0],
                    torch.tensor([e], dtype=torch.float64, device=cuda))
    for chunk in (fw[:5], fw[5:]):
        got = heat_from_edge_mean(base, chunk).host()
        want = [heat_parameters_from_edge_mean(e, f) for f in chunk]
        assert got == want


def test_type5_sheets_batch_equals_per_side(small, cuda):
    """Sides of one topology with their own geometry, smoothing counts and sheet
    stencils in one csg_type5_sheets_batch call equal per-side csg_type5_sheets."""

    import dataclasses

    from cat_surface_gpu.surface_stencil import (edge_means, heat_parameters_from_edge_mean,
                                                 smooth_geometry_many, type5_sheets_many)

    g, dev = small
    surf = torch.as_tensor(g["surface"], device=cuda)
    flipped = dataclasses.replace(
        dev, sheet_indices=torch.flip(dev.sheet_indices, (0,)).contiguous(),
        sheet_weights=torch.flip(dev.sheet_weights, (0,)).contiguous())
    assert flipped.topology is dev.topology
    gen = torch.Generator(device="cpu").manual_seed(7)
    noise = (torch.rand(surf.shape, generator=gen) - 0.5).to(cuda) * 0.05
    sides = [dev, flipped, dev, flipped]
    geoms = [surf, surf * 1.25, surf + noise, surf * 0.8 - noise]
    fw = (30.0, 12.0)  # 16 and 3 heat steps on the unscaled surface
    ems = edge_means(sides, geoms)
    gis = [heat_parameters_from_edge_mean(em, 50.0)[0] for em in ems]
    assert len(set(heat_parameters_from_edge_mean(em, fw[0])[0] for em in ems)) > 1
    smoothed = smooth_geometry_many(sides, geoms, gis)
    for dt in (torch.float64, torch.float32):
        batched = type5_sheets_many(sides, geoms, fw, dt)
        for st, geo, em, gi, sm, got in zip(sides, geoms, ems, gis, smoothed, batched):
            params = [heat_parameters_from_edge_mean(em, f) for f in fw]
            single = st._type5(geo, gi, params, dt, smoothed=sm.contiguous())
            assert got.dtype == dt and torch.equal(got, single)
    # more rows than one loop holds: split into calls of at most 32 rows
    many = type5_sheets_many([dev] * 17, [surf] * 17, fw)
    for got in many:
        assert torch.equal(got, many[0])


def _feature_stencil(g, cuda):
    """The reference fixture's stencil (642-vertex sphere, resampling weights) on the GPU."""

    import cat_surface_gpu as csg

    n = g["verts"].shape[0]
    return csg.SurfaceStencil(
        sphere_points=g["verts"], faces=g["faces"], surface_indices=g["surf_idx"],
        surface_weights=g["surf_w"], sheet_indices=np.zeros((1, 1, 3), np.int32),
        sheet_weights=np.zeros((1, 1, 3)), nx=1, ny=1, source_points=n).to(cuda)


def test_rotation_feature_matches_golden(golden, cuda):
    import cat_surface_gpu as csg
    from cat_surface_gpu.rotation_feature import compute_rotation_feature

    g = golden("rotation_feature")
    n = g["verts"].shape[0]
    st = csg.SurfaceStencil(
        sphere_points=g["verts"], faces=g["faces"], surface_indices=g["surf_idx"],
        surface_weights=g["surf_w"], sheet_indices=np.zeros((1, 1, 3), np.int32),
        sheet_weights=np.zeros((1, 1, 3)), nx=1, ny=1, source_points=n).to(cuda)
    res = compute_rotation_feature(st, torch.as_tensor(g["surface"], device=cuda),
                                   heat_fwhm=15.0)
    assert res.relative_residual < 1e-9
    np.testing.assert_allclose(_np(res.values), g["feature"], atol=1e-6)


def test_depth_potential_batch_matches_single(golden, cuda):
    """Both sides solved in one persistent launch give exactly the per-side values."""

    import cat_surface_gpu as csg
    from cat_surface_gpu.rotation_feature import (compute_rotation_feature,
                                                  compute_rotation_features, depth_potential)

    g = golden("rotation_feature")
    n = g["verts"].shape[0]
    st = csg.SurfaceStencil(
        sphere_points=g["verts"], faces=g["faces"], surface_indices=g["surf_idx"],
        surface_weights=g["surf_w"], sheet_indices=np.zeros((1, 1, 3), np.int32),
        sheet_weights=np.zeros((1, 1, 3)), nx=1, ny=1, source_points=n).to(cuda)
    a = torch.as_tensor(g["smoothed"], dtype=torch.float32, device=cuda)
    rng = np.random.default_rng(3)
    b = a * 1.1 + torch.as_tensor(rng.normal(0.0, 0.01, a.shape), dtype=torch.float32,
                                  device=cuda)
    xa, _, ra = depth_potential(st, a)
    xb, _, rb = depth_potential(st, b)
    xs, _, rs = depth_potential(st, torch.stack([a, b]))
    assert torch.equal(xs[0], xa) and torch.equal(xs[1], xb)
    assert float(rs[0]) == float(ra) and float(rs[1]) == float(rb)
    np.testing.assert_allclose(_np(xa), g["depth"], atol=1e-8 * np.abs(g["depth"]).max())
    surf = torch.as_tensor(g["surface"], device=cuda)
    pair = compute_rotation_features((st, st), (surf, surf), heat_fwhms=(15.0, 10.0))
    solo = compute_rotation_feature(st, surf, heat_fwhm=10.0)
    assert torch.equal(pair[1].values, solo.values)
    np.testing.assert_allclose(_np(pair[0].values), g["feature"], atol=1e-6)


def test_pcg_depth_potential_matches_reference_and_oracle(golden, cuda):
    """The PCG solver against the reference's CG fixture (642 vertices) and the
    oracle's CG; the feature built on it."""

    from cat_surface_gpu.rotation_feature import compute_rotation_feature, depth_potential
    from oracle import pipeline as P

    g = golden("rotation_feature")
    st = _feature_stencil(g, cuda)
    x, its, rel = depth_potential(st, torch.as_tensor(g["smoothed"], device=cuda), solver="pcg",
                                  tolerance=1e-9)
    scale = np.abs(g["depth_cg"]).max()
    assert its == int(g["cg_iterations"])
    assert float(rel) <= 1e-9
    np.testing.assert_allclose(_np(x), g["depth_cg"], rtol=0, atol=1e-9 * scale)
    ref = P.depth_potential(g["smoothed"], g["faces"], solver="pcg", tolerance=1e-9)
    np.testing.assert_allclose(_np(x), ref.values, rtol=0, atol=1e-9 * scale)
    res = compute_rotation_feature(st, torch.as_tensor(g["surface"], device=cuda),
                                   heat_fwhm=15.0, solver="pcg")
    np.testing.assert_allclose(_np(res.values), g["feature_cg"], atol=1e-6)
    assert res.iterations == int(g["cg_iterations"])


@pytest.mark.parametrize("batch", [1, 4, 9])
def test_pcg_depth_potential_at_40962_vertices(batch, cuda):
    """PCG on icosphere-6 geometries (40,962 vertices): every mesh of a batch
    equals the same mesh solved alone bit for bit (fixed reduction order), it
    reaches the tolerance, and it agrees with the oracle's CG (measured: 495
    iterations, as the oracle); when the workgroups of a batch all claim one XCD
    (as on a partition reporting one XCC_ID) the meshes of the other XCDs get no
    group, the solve notices and the one-workgroup-per-mesh fallback returns the
    same values bit for bit (the reductions follow fixed row chunks)."""

    import cat_surface_gpu as csg
    from cat_surface_gpu import _lib
    from cat_surface_gpu.rotation_feature import depth_potential
    from cat_surface_gpu.stencil_builder import icosphere
    from oracle import pipeline as P

    v, f = icosphere(6)
    n = v.shape[0]
    idx = np.repeat(np.arange(n, dtype=np.int32)[:, None], 3, axis=1)
    w = np.tile(np.array([1.0, 0.0, 0.0]), (n, 1))
    st = csg.SurfaceStencil(
        sphere_points=v, faces=f.astype(np.int32), surface_indices=idx, surface_weights=w,
        sheet_indices=np.zeros((1, 1, 3), np.int32), sheet_weights=np.zeros((1, 1, 3)),
        nx=1, ny=1, source_points=n).to(cuda)
    rng = np.random.default_rng(40 + batch)
    # a smooth bumpy surface at millimetre scale, as the pipeline's fwhm-15 geometry
    dirs = rng.normal(size=(6, 3))
    bump = 1.0 + 0.03 * np.sin(v @ dirs.T * 3.0).sum(1)
    geoms = np.stack([v * 60.0 * (bump * (1.0 + 0.01 * k))[:, None] for k in range(batch)])
    pts = torch.as_tensor(geoms, dtype=torch.float32, device=cuda)
    x, its, rel = depth_potential(st, pts, solver="pcg", tolerance=1e-9)
    x1, its1, rel1 = depth_potential(st, pts[-1], solver="pcg", tolerance=1e-9)
    assert torch.equal(x[-1], x1) and float(rel[-1]) == float(rel1)
    assert float(rel.max()) <= 1e-9
    ref = P.depth_potential(geoms[-1].astype(np.float32), f, solver="pcg", tolerance=1e-9)
    scale = np.abs(ref.values).max()
    np.testing.assert_allclose(_np(x1), ref.values, rtol=0, atol=1e-7 * scale)
    print(f"\nPCG 40,962 vertices, batch {batch}: {its1} iterations (oracle {ref.iterations}), "
          f"relres {float(rel1):.2e}")
    if batch == 1:
        return  # one mesh: forcing every claim onto XCD 0 leaves no group short
    lib = _lib.lib()
    lib.csg_reset_sor_xcd_check()
    lib.csg_set_sor_force_xcd(0)
    try:
        xf, _, relf = depth_potential(st, pts, solver="pcg", tolerance=1e-9)
    finally:
        lib.csg_set_sor_force_xcd(-1)
        lib.csg_reset_sor_xcd_check()
    # the chunked reductions make the fallback's iterates the grouped ones exactly
    assert torch.equal(xf, x) and torch.equal(relf, rel)


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("level,batch", [(3, 1), (6, 1), (6, 4), (5, 9)])
def test_sor_xcd_local_equals_grid_barrier(level, batch, mode, cuda):
    """Each mesh's sweeps on one XCD give exactly the device-wide-barrier values."""

    import cat_surface_gpu as csg
    from cat_surface_gpu.rotation_feature import (depth_potential, set_sor_patches,
                                                  set_sor_placement)
    from cat_surface_gpu.stencil_builder import icosphere

    v, f = icosphere(level)
    n = v.shape[0]
    idx = np.repeat(np.arange(n, dtype=np.int32)[:, None], 3, axis=1)
    w = np.tile(np.array([1.0, 0.0, 0.0]), (n, 1))
    st = csg.SurfaceStencil(
        sphere_points=v, faces=f.astype(np.int32), surface_indices=idx, surface_weights=w,
        sheet_indices=np.zeros((1, 1, 3), np.int32), sheet_weights=np.zeros((1, 1, 3)),
        nx=1, ny=1, source_points=n).to(cuda)
    rng = np.random.default_rng(level * 100 + batch)
    bumps = 1.0 + 0.05 * rng.standard_normal((batch, n, 1))
    pts = torch.as_tensor(v[None] * 60.0 * bumps, dtype=torch.float32, device=cuda)
    prev = set_sor_placement(mode)
    prev_patches = set_sor_patches(False)
    try:
        xl, _, rl = depth_potential(st, pts, max_iter=200)
        set_sor_placement(0)
        xg, _, rg = depth_potential(st, pts, max_iter=200)
    finally:
        set_sor_placement(prev)
        set_sor_patches(prev_patches)
    assert torch.isfinite(xl).all()
    assert torch.equal(xl, xg)
    assert torch.equal(rl, rg)


def _timed_launches(name: str) -> int:
    """Launches of a timed kernel since the last read (reading clears them)."""

    import ctypes

    from cat_surface_gpu import _lib

    lib = _lib.lib()
    k = 0
    while True:
        label = lib.csg_kernel_timing_name(k)
        if not label:
            return 0
        if label.decode() == name:
            ms, nbytes, launches = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
            _lib.check(lib.csg_kernel_timing_read(k, ctypes.byref(ms), ctypes.byref(nbytes),
                                                  ctypes.byref(launches)), "timing")
            return int(launches.value)
        k += 1


@pytest.mark.parametrize("level,batch,iters,launches", [
    (6, 1, 300, 1), (6, 4, 300, 1), (6, 8, 300, 1), (4, 2, 301, 1),
    (5, 3, 1, 1), (5, 9, 300, 2), (6, 12, 301, 2), (5, 17, 300, 3)])
def test_sor_patched_equals_per_colour(level, batch, iters, launches, cuda):
    """The SOR blocked in time on mesh patches (one exchange per sweep) gives
    exactly the values of one synchronisation per colour.  One XCD holds one
    mesh's 32 patches at once, so batches beyond 8 meshes run as consecutive
    launches of 8 (``launches``), with the same values."""

    import cat_surface_gpu as csg
    from cat_surface_gpu import _lib
    from cat_surface_gpu.rotation_feature import depth_potential, set_sor_patches
    from cat_surface_gpu.stencil_builder import icosphere

    v, f = icosphere(level)
    n = v.shape[0]
    idx = np.repeat(np.arange(n, dtype=np.int32)[:, None], 3, axis=1)
    w = np.tile(np.array([1.0, 0.0, 0.0]), (n, 1))
    st = csg.SurfaceStencil(
        sphere_points=v, faces=f.astype(np.int32), surface_indices=idx, surface_weights=w,
        sheet_indices=np.zeros((1, 1, 3), np.int32), sheet_weights=np.zeros((1, 1, 3)),
        nx=1, ny=1, source_points=n).to(cuda)
    rng = np.random.default_rng(level * 10 + batch)
    bumps = 1.0 + 0.05 * rng.standard_normal((batch, n, 1))
    pts = torch.as_tensor(v[None] * 60.0 * bumps, dtype=torch.float32, device=cuda)
    prev = set_sor_patches(False)
    try:
        xg, _, rg = depth_potential(st, pts, max_iter=iters)
        set_sor_patches(True)
        _lib.lib().csg_kernel_timing(1)
        _timed_launches("k_sor_patched")  # reading clears the totals
        xp, _, rp = depth_potential(st, pts, max_iter=iters)
        used = _timed_launches("k_sor_patched")
        _lib.lib().csg_kernel_timing(0)
    finally:
        set_sor_patches(prev)
    assert torch.isfinite(xp).all()
    assert torch.equal(xp, xg)
    assert torch.equal(rp, rg)
    assert used == launches


@pytest.mark.parametrize("patched", [True, False])
def test_sor_partition_falls_back_to_grid_barrier(patched, cuda):
    """A device whose workgroups all report one XCC_ID (a CPX/DPX partition,
    forced from the host) leaves the XCD groups of meshes 1..3 without
    workgroups: the solve notices, repeats on the device-wide barrier and
    returns exactly its values, and the device keeps to that placement."""

    import cat_surface_gpu as csg
    from cat_surface_gpu import _lib
    from cat_surface_gpu.rotation_feature import (depth_potential, set_sor_patches,
                                                  set_sor_placement)
    from cat_surface_gpu.stencil_builder import icosphere

    v, f = icosphere(6)
    n = v.shape[0]
    idx = np.repeat(np.arange(n, dtype=np.int32)[:, None], 3, axis=1)
    w = np.tile(np.array([1.0, 0.0, 0.0]), (n, 1))
    st = csg.SurfaceStencil(
        sphere_points=v, faces=f.astype(np.int32), surface_indices=idx, surface_weights=w,
        sheet_indices=np.zeros((1, 1, 3), np.int32), sheet_weights=np.zeros((1, 1, 3)),
        nx=1, ny=1, source_points=n).to(cuda)
    rng = np.random.default_rng(17)
    bumps = 1.0 + 0.05 * rng.standard_normal((4, n, 1))
    pts = torch.as_tensor(v[None] * 60.0 * bumps, dtype=torch.float32, device=cuda)
    lib = _lib.lib()
    lib.csg_reset_sor_xcd_check()
    prev_place = set_sor_placement(0)
    prev_patches = set_sor_patches(False)
    try:
        xg, _, rg = depth_potential(st, pts, max_iter=50)
        set_sor_placement(2)
        set_sor_patches(patched)
        xl, _, _ = depth_potential(st, pts, max_iter=50)  # the real placement: no fallback
        assert lib.csg_reset_sor_xcd_check() == 0
        lib.csg_set_sor_force_xcd(0)
        xf, _, rf = depth_potential(st, pts, max_iter=50)
        assert lib.csg_set_sor_force_xcd(-1) == 0
        assert lib.csg_reset_sor_xcd_check() == 1
    finally:
        lib.csg_set_sor_force_xcd(-1)
        lib.csg_reset_sor_xcd_check()
        set_sor_placement(prev_place)
        set_sor_patches(prev_patches)
    assert torch.equal(xl, xg)
    assert torch.equal(xf, xg)
    assert torch.equal(rf, rg)


def test_rotation_index_costs_and_search(golden, cuda):
    import cat_surface_gpu as csg

    g = golden("rotation")
    idx = csg.RotationGridIndex.from_geometry(g["verts"], g["faces"], grid_size=32, margin=1,
                                              device=cuda)
    np.testing.assert_array_equal(idx.candidate_offsets, g["cand_offsets"])
    np.testing.assert_array_equal(idx.candidate_faces, g["cand_faces"])
    dev = idx.to(cuda)
    inputs = dev.prepare_cost_inputs(g["src_pts"], g["src_vals"], g["tgt_vals"])
    costs = _np(dev.cost_batch_prepared(inputs, torch.as_tensor(g["seeds"])))
    # float32 location (barycentric weights in another operation order than torch's) and
    # float64 sums in another order: measured 3.4e-7 relative on MI355X (3.8e-7 against
    # the oracle, weights 1.1e-6 apart); the faces themselves agree exactly
    print(f"\nseed costs: max rel vs reference {np.abs(costs / g['seed_costs'] - 1).max():.2e}")
    np.testing.assert_allclose(costs, g["seed_costs"], rtol=1e-6)
    faces, bary = dev.locate(inputs, torch.as_tensor(g["seeds"][:5]))
    # point location is index work: the located faces are exact
    np.testing.assert_array_equal(_np(faces)[:, :200], g["loc_faces"])
    ocost, oface, obary = O.rotation_costs(g["src_pts"], g["src_vals"], g["tgt_vals"],
                                           g["verts"], g["faces"], g["seeds"], 32, 1,
                                           return_location=True)
    np.testing.assert_array_equal(_np(faces), oface[:5])
    print(f"located weights: max |GPU - oracle| {np.abs(_np(bary) - obary[:5]).max():.2e}; "
          f"costs: max rel {np.abs(costs / ocost - 1).max():.2e}")
    np.testing.assert_allclose(_np(bary), obary[:5], rtol=0, atol=2e-6)
    np.testing.assert_allclose(costs, ocost, rtol=1e-6)
    angle, cost, seeds = dev.coarse_seed_search_prepared(inputs)
    np.testing.assert_array_equal(_np(angle), g["seed_angle"])
    a, c, it = dev.refine_nelder_mead_prepared(inputs, angle, max_iter=60)
    print(f"Nelder-Mead: |angle - reference| {np.abs(_np(a) - g['nm_angle']).max():.2e}, "
          f"iterations {int(it)} vs {int(g['nm_iters'])}, cost rel "
          f"{abs(float(c) / float(g['nm_cost']) - 1):.2e}")
    # the simplex moves are decided by cost comparisons that no reduction-order
    # difference flips here: the trajectory, its end point and length are the reference's
    np.testing.assert_allclose(_np(a), g["nm_angle"], rtol=0, atol=1e-12)
    assert int(it) == int(g["nm_iters"])
    np.testing.assert_allclose(float(c), float(g["nm_cost"]), rtol=2e-6)


def test_rotation_costs_match_oracle_on_random_angles(golden, cuda):
    import cat_surface_gpu as csg

    g = golden("rotation")
    rng = np.random.default_rng(8)
    angles = rng.uniform(-0.7, 0.7, size=(64, 3))
    idx = csg.RotationGridIndex.from_geometry(g["verts"], g["faces"], grid_size=32, margin=1,
                                              device=cuda).to(cuda)
    costs = _np(idx.cost_batch(g["src_pts"], g["src_vals"], g["tgt_vals"],
                               torch.as_tensor(angles)))
    ref = O.rotation_costs(g["src_pts"], g["src_vals"], g["tgt_vals"], g["verts"], g["faces"],
                           angles, 32, 1)
    np.testing.assert_allclose(costs, ref, rtol=1e-6)


def test_identity_rotation_cost_is_zero(cuda):
    """Vertex values sampled at the vertices themselves cost nothing (reference test)."""

    import cat_surface_gpu as csg
    from cat_surface_gpu.stencil_builder import icosphere

    v, f = icosphere(2)
    vals = np.arange(v.shape[0], dtype=np.float64)
    idx = csg.RotationGridIndex.from_geometry(v, f, grid_size=16, margin=2, device=cuda).to(cuda)
    cost = _np(idx.cost_batch(v.astype(np.float32), vals, vals, torch.zeros((1, 3))))
    assert cost[0] < 1e-6, cost


def test_warp_matches_golden(golden, cuda):
    import cat_surface_gpu as csg

    g = golden("warp")
    pts = torch.as_tensor(g["points"], device=cuda)
    flow = torch.as_tensor(g["flow"], device=cuda)
    np.testing.assert_allclose(_np(csg.apply_flow_to_sphere(pts, flow, inverse=True,
                                                            device=cuda)),
                               g["warp_inverse"], rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(_np(csg.apply_flow_to_sphere(pts, flow, inverse=False,
                                                            device=cuda)),
                               g["warp_forward"], rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(_np(csg.apply_flow_to_sphere(pts * 3.0, flow,
                                                            unit_sphere_vertices=pts,
                                                            device=cuda)),
                               g["warp_cached"], rtol=1e-12, atol=1e-13)


def test_average_xz_matches_golden(golden, cuda):
    from cat_surface_gpu import _lib

    g = golden("warp")
    x = torch.as_tensor(g["avg_x"], dtype=torch.float64, device=cuda)
    z = torch.as_tensor(g["avg_z"], dtype=torch.float64, device=cuda)
    out = torch.empty(x.shape, dtype=torch.float32, device=cuda)
    _lib.check(_lib.lib().csg_average_xz(_lib.ptr(x), _lib.ptr(z), int(x.shape[0]), 1,
                                         _lib.doubles(np.eye(3).reshape(-1)), _lib.ptr(out),
                                         _lib.stream_handle(cuda)), "avg")
    np.testing.assert_array_equal(_np(out), g["avg_out"])


def _ulps(a, b):
    """Distance in float32 units in the last place."""

    ia = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    ib = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, np.int64(-2**31) - ia, ia)
    ib = np.where(ib < 0, np.int64(-2**31) - ib, ib)
    return np.abs(ia - ib)


def test_prepare_sphere_matches_golden_and_oracle(golden, cuda):
    """csg_prepare_sphere: the reference's sequential-centroid normalisation and
    float64 rotation (golden vectors), and the same on a 61,442-vertex sphere
    against the oracle.  The centroid is the same strictly sequential sum, so the
    normalisation is bit-exact; the rotation's 3-term float64 dot products may
    round differently from numpy's matmul, which at most moves the float32
    result by one unit in the last place."""

    from cat_surface_gpu.gpu_pipeline import _check_normalised, _prepare_sphere, _sphere_centroids

    g = golden("warp")
    bad = torch.zeros(1, dtype=torch.int32, device=cuda)
    v = torch.as_tensor(g["norm_in"], dtype=torch.float32, device=cuda)
    c = _sphere_centroids([v])[0]
    norm = _prepare_sphere(v, centroid=c, bad=bad)
    np.testing.assert_array_equal(_np(norm), g["norm_out"])
    rot = _prepare_sphere(norm, angles=g["rot_angles"][1])
    assert _ulps(_np(rot), g["rotv_out"]).max() <= 1
    both = _prepare_sphere(v, centroid=c, angles=g["rot_angles"][1], bad=bad)
    np.testing.assert_array_equal(_np(both), _np(rot))
    _check_normalised(bad)

    # spheres of several sizes in one centroid launch (chunk tails, more than one
    # launch's worth of spheres), each against the oracle's sequential sum
    rng = np.random.default_rng(11)
    sizes = [61442, 1, 255, 256, 257, 1000] + [int(k) for k in rng.integers(2, 3000, 64)]
    spheres = []
    for n in sizes:
        p = rng.normal(size=(n, 3))
        p = p / np.linalg.norm(p, axis=1, keepdims=True) * 100.0 + rng.normal(size=3) * 10
        spheres.append(p.astype(np.float32))
    dev = [torch.as_tensor(p, device=cuda) for p in spheres]
    cs = _np(_sphere_centroids(dev))
    for p, got in zip(spheres, cs):
        np.testing.assert_array_equal(got, np.cumsum(p.astype(np.float64), axis=0)[-1]
                                      / p.shape[0])
    big, dv = spheres[0], dev[0]
    angles = np.array([0.7, -1.1, 2.3])
    got = _prepare_sphere(dv, centroid=_sphere_centroids([dv])[0], bad=bad)
    np.testing.assert_array_equal(_np(got), O.normalise_sphere(big))
    got = _prepare_sphere(dv, centroid=_sphere_centroids([dv])[0], angles=angles, bad=bad)
    ulp = _ulps(_np(got), O.rotate_vertices(O.normalise_sphere(big), angles))
    assert ulp.max() <= 1 and (ulp > 0).mean() < 1e-3
    _check_normalised(bad)

    # a vertex at the centroid cannot be normalised (the reference raises ValueError)
    degenerate = torch.as_tensor(np.array([[1, 0, 0], [-1, 0, 0], [0, 0, 0]], np.float32),
                                 device=cuda)
    _prepare_sphere(degenerate, centroid=_sphere_centroids([degenerate])[0], bad=bad)
    with pytest.raises(ValueError, match="cannot be normalized"):
        _check_normalised(bad)
    with pytest.raises(ValueError, match="cannot be normalized"):
        O.normalise_sphere(_np(degenerate))


def test_grid_index_build_equals_two_phase(cuda):
    """csg_grid_index_build (one asynchronous call, bounded pair capacity) lists the
    same cells as the two-phase plan/fill index when the capacity suffices, and
    reports the required count (staying in bounds) when it does not."""

    from cat_surface_gpu.locate import build_grid_index, build_grid_index_async
    from cat_surface_gpu.stencil_builder import icosphere

    sv, sf = icosphere(4)
    rng = np.random.default_rng(5)
    pts = torch.as_tensor(0.9 * sv + rng.normal(scale=0.01, size=sv.shape), device=cuda)
    faces = torch.as_tensor(sf, device=cuda)
    for grid, margin in ((32, 1), (16, 2)):
        ref = build_grid_index(pts, faces, grid, margin)
        n = int(ref.cell_faces.numel())
        for capacity in (n, n + 777):
            req = torch.zeros(1, dtype=torch.int32, device=cuda)
            got = build_grid_index_async(pts, faces, grid, margin, capacity, req)
            assert int(req.item()) == n
            np.testing.assert_array_equal(_np(got.cell_off), _np(ref.cell_off))
            np.testing.assert_array_equal(_np(got.cell_faces)[:n], _np(ref.cell_faces))
        req = torch.zeros(1, dtype=torch.int32, device=cuda)
        small = build_grid_index_async(pts, faces, grid, margin, n // 3, req)
        assert int(req.item()) == n
        off = _np(small.cell_off)
        assert off[-1] == n // 3 and np.all(np.diff(off) >= 0)
        listed = _np(small.cell_faces)
        assert listed.min() >= 0 and listed.max() < sf.shape[0]


def test_deferred_stencil_equals_synchronous(cuda):
    """A stencil built without host round trips (deferred checks) equals the one
    built with them, and its checks pass."""

    from cat_surface_gpu.locate import check_all, closest_triangles
    from cat_surface_gpu.stencil_builder import build_device_stencil, icosphere

    sv, sf = icosphere(5)
    rng = np.random.default_rng(9)
    sphere = torch.as_tensor((sv + rng.normal(scale=0.003, size=sv.shape)).astype(np.float32),
                             device=cuda)
    faces = torch.as_tensor(sf, device=cuda)
    ref = build_device_stencil(sphere, faces, nx=64, ny=32, device=cuda)  # learns capacities
    got, checks = build_device_stencil(sphere, faces, nx=64, ny=32, device=cuda, deferred=True)
    assert all(c is not None for c in checks)
    assert check_all(checks) == [True, True]
    for name in ("sphere_points", "surface_indices", "surface_weights", "sheet_indices",
                 "sheet_weights", "unit_sphere_points"):
        np.testing.assert_array_equal(_np(getattr(got, name)), _np(getattr(ref, name)), name)
    # a query set whose closest triangles lie beyond one cell fails its check
    far_q = torch.as_tensor(np.array([[0.0, 0.0, 0.0], [0.05, 0.0, 0.0]]), device=cuda)
    from cat_surface_gpu.stencil_builder import _GRID_SURFACE

    face, idx, w, check = closest_triangles(far_q, sphere.double(), faces,
                                            grid_size=_GRID_SURFACE, margin=1, deferred=True)
    assert check is not None and check_all([check]) == [False]


def test_stencil_builder_matches_brute_force(cuda):
    """Closest-triangle stencil entries equal an exhaustive CPU search."""

    from cat_surface_gpu.stencil_builder import build_stencil, icosphere, uv_points

    sv, sf = icosphere(3)
    rng = np.random.default_rng(3)
    sphere = (sv + rng.normal(scale=0.01, size=sv.shape)).astype(np.float32)
    st = build_stencil(sphere, sf, nx=32, ny=16, n_triangles=1280, device=cuda)
    # the same normalisation the builder applies
    pts = sphere.astype(np.float64)
    pts = pts - pts.mean(0)
    pts /= np.linalg.norm(pts, axis=1, keepdims=True)
    radius = 0.975 * np.linalg.norm(pts, axis=1).mean()
    ico, _ = icosphere(3)
    out_sphere = (pts.min(0) + pts.max(0))[None] + radius * ico
    face, w = O.closest_triangles(pts, sf, out_sphere)
    np.testing.assert_array_equal(st.surface_indices, sf[face])
    np.testing.assert_allclose(st.surface_weights, w, atol=1e-12)
    assert np.all(st.surface_weights > -1e-12)
    np.testing.assert_allclose(st.surface_weights.sum(1), 1.0, atol=1e-12)
    unit = st.sphere_points / np.linalg.norm(st.sphere_points, axis=1, keepdims=True)
    face2, w2 = O.closest_triangles(unit, st.faces, uv_points(32, 16))
    np.testing.assert_array_equal(st.sheet_indices.reshape(-1, 3), st.faces[face2])
    np.testing.assert_allclose(st.sheet_weights.reshape(-1, 3), w2, atol=1e-12)
    np.testing.assert_allclose(np.linalg.norm(st.unit_sphere_points, axis=1), 1.0, atol=1e-6)


def _smooth_rotation_problem(cuda, angle=(0.2, -0.1, 0.15), tgt_level=4, src_level=3):
    """Target icosphere carrying a smooth field; source points whose values are
    the field at their image under ``angle``'s rotation."""

    import cat_surface_gpu as csg
    from cat_surface_gpu.stencil_builder import icosphere
    from cat_surface_gpu.surface_rotation import rotation_matrix_np

    def field(p):
        return np.sin(2.0 * p[:, 0]) + np.cos(3.0 * p[:, 1]) * p[:, 2] + 0.5 * p[:, 0] * p[:, 1]

    v, f = icosphere(tgt_level)
    src, _ = icosphere(src_level)
    m = rotation_matrix_np(angle)
    idx = csg.RotationGridIndex.from_geometry(v, f, grid_size=32, margin=1,
                                              device=cuda).to(cuda)
    inputs = idx.prepare_cost_inputs(src.astype(np.float32), field(src @ m.T), field(v))
    return idx, inputs, m


def test_rotation_cost_gradient_matches_finite_differences(cuda):
    from cat_surface_gpu.surface_rotation import angles_from_matrix, axis_angle_matrix

    idx, inputs, m = _smooth_rotation_problem(cuda)
    m0 = axis_angle_matrix([0.05, -0.03, 0.04]) @ m
    cost, grad, hess = idx.cost_matrices(inputs, m0[None], gradient=True)
    cost, grad, hess = float(cost[0]), _np(grad[0]), _np(hess[0])
    # the matrix-input cost is the angle-input cost of the same rotation
    ref = float(idx.cost_batch_prepared(inputs, torch.as_tensor(angles_from_matrix(m0)[None]))[0])
    assert abs(cost - ref) <= 1e-4 * ref
    assert float(idx.cost_matrices(inputs, m0[None])[0]) == cost
    h = 1e-3
    fd = np.empty(3)
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        plus = float(idx.cost_matrices(inputs, (axis_angle_matrix(e) @ m0)[None])[0])
        minus = float(idx.cost_matrices(inputs, (axis_angle_matrix(-e) @ m0)[None])[0])
        fd[k] = (plus - minus) / (2.0 * h)
    np.testing.assert_allclose(grad, fd, rtol=0, atol=2e-2 * np.abs(fd).max())
    # Gauss-Newton matrix: symmetric positive semi-definite
    np.testing.assert_allclose(hess, hess.T)
    assert np.linalg.eigvalsh(hess).min() >= -1e-9 * np.abs(hess).max()


def _field_maps(m, cuda, nx=512, ny=256):
    """Equirectangular maps of the smooth field: target T(x) = f(x), source
    S(x) = f(M x), so that S(Mᵀ x) = T(x)."""

    from cat_surface_gpu.stencil_builder import uv_points

    def field(p):
        return np.sin(2.0 * p[:, 0]) + np.cos(3.0 * p[:, 1]) * p[:, 2] + 0.5 * p[:, 0] * p[:, 1]

    x = uv_points(nx, ny)
    to = dict(dtype=torch.float64, device=cuda)
    return (torch.as_tensor(field(x @ m.T).reshape(ny, nx), **to),
            torch.as_tensor(field(x).reshape(ny, nx), **to))


def test_map_rotation_cost_is_minimal_at_the_true_rotation(cuda):
    from cat_surface_gpu.surface_rotation import axis_angle_matrix, map_costs, rotation_matrix_np

    m = rotation_matrix_np((0.2, -0.1, 0.15))
    smap, tmap = _field_maps(m, cuda)
    others = [axis_angle_matrix(d) @ m for d in ([0.05, 0, 0], [0, -0.05, 0], [0, 0, 0.05])]
    mats = torch.as_tensor(np.stack([m, np.eye(3)] + others), device=cuda)
    costs = _np(map_costs(smap, tmap, mats))
    assert costs[0] < 1e-3 * costs[1:].min()
    # the stride only subsamples the target pixels
    full = _np(map_costs(smap, tmap, mats, stride=1))
    assert int(np.argmin(full)) == 0


@pytest.mark.parametrize("candidates", ["seeds", "hopf", "hopf-maps"])
def test_adjoint_rotation_search_reaches_nelder_mead_cost(cuda, candidates):
    """Top candidates refined by gradient steps cost no more than the reference's
    seed + Nelder-Mead choice (within 1e-6 relative) and find the true rotation."""

    from cat_surface_gpu.surface_rotation import search_rotations

    idx, inputs, m = _smooth_rotation_problem(cuda)
    nm = search_rotations([(idx, inputs)])[0]
    problem = (idx, inputs)
    if candidates == "hopf-maps":
        problem = (idx, inputs, _field_maps(m, cuda))
        candidates = "hopf"
    adj = search_rotations([problem], method="adjoint", candidates=candidates)[0]
    assert float(adj.cost) <= float(nm.cost) * (1.0 + 1e-6)
    rel = _np(adj.rotation_matrix).astype(np.float64) @ m.T
    assert np.arccos(np.clip((np.trace(rel) - 1.0) / 2.0, -1.0, 1.0)) < 5e-3


def test_adjoint_rotation_search_on_reference_problem(golden, cuda):
    import cat_surface_gpu as csg
    from cat_surface_gpu.surface_rotation import search_rotations

    g = golden("rotation")
    idx = csg.RotationGridIndex.from_geometry(g["verts"], g["faces"], grid_size=32, margin=1,
                                              device=cuda).to(cuda)
    inputs = idx.prepare_cost_inputs(g["src_pts"], g["src_vals"], g["tgt_vals"])
    adj = search_rotations([(idx, inputs)], method="adjoint")[0]
    assert float(adj.cost) <= float(g["nm_cost"]) * (1.0 + 1e-6)
    np.testing.assert_allclose(_np(adj.seed_angle), g["seed_angle"])


@pytest.mark.parametrize("level,iterations", [(6, (50, 13, 0, 7)), (3, (5, 2)), (5, tuple(range(9)))])
def test_geometry_smoothing_padded_equals_packed(level, iterations, cuda):
    """The padded float4 schedule (full workspace) gives the bits of the packed one."""

    from cat_surface_gpu import _lib
    from cat_surface_gpu.stencil_builder import icosphere
    from cat_surface_gpu.surface_stencil import topology_for

    v, f = icosphere(level)
    n = v.shape[0]
    topo = topology_for(f.astype(np.int32), n, cuda)
    rng = np.random.default_rng(level)
    batch = len(iterations)
    geom = torch.as_tensor(v[None] * 50.0 * (1.0 + 0.05 * rng.standard_normal((batch, n, 1))),
                           dtype=torch.float32, device=cuda).contiguous()
    its = (_lib.ctypes.c_int * batch)(*iterations)
    lib = _lib.lib()
    outs = []
    for nbytes in (lib.csg_smooth_geometry_batch_workspace_bytes(n, batch), 3 * n * batch * 4):
        ws = _lib.workspace(nbytes, cuda)
        out = torch.empty_like(geom)
        _lib.check(lib.csg_smooth_geometry_batch(
            _lib.ptr(geom), batch, _lib.ptr(topo.nbr_off), _lib.ptr(topo.nbr), n, its,
            _lib.ptr(out), _lib.ptr(ws), nbytes, _lib.stream_handle(cuda)),
            "csg_smooth_geometry_batch")
        outs.append(out)
    torch.cuda.synchronize()
    assert torch.equal(outs[0], outs[1])
    for b, k in enumerate(iterations):
        if k == 0:
            assert torch.equal(outs[0][b], geom[b])


@pytest.mark.parametrize("level,iterations,patches", [
    (6, (50, 13, 0, 7), (256, 8)), (6, (283, 291, 17, 300), (512, 4)),
    (3, (5, 2), (64, 3)), (5, tuple(range(9)), (256, 8)), (4, (20, 21), (100000, 6))])
def test_geometry_smoothing_patched_equals_per_iteration(level, iterations, patches, cuda):
    """Patch-local iterations in LDS (csg_smooth_geometry_patched) give the bits of
    one launch per iteration (csg_smooth_geometry_batch), for iteration counts
    below, at and between multiples of the rings per launch, and for one patch
    covering the whole mesh; both equal the oracle's float64 arithmetic
    (correctly rounded divisions, the same neighbour order) bit for bit."""

    from cat_surface_gpu import _lib
    from cat_surface_gpu.stencil_builder import icosphere
    from cat_surface_gpu.surface_stencil import smoothing_patches, topology_for

    v, f = icosphere(level)
    n = v.shape[0]
    topo = topology_for(f.astype(np.int32), n, cuda)
    rng = np.random.default_rng(level + 100)
    batch = len(iterations)
    geom = torch.as_tensor(v[None] * 50.0 * (1.0 + 0.05 * rng.standard_normal((batch, n, 1))),
                           dtype=torch.float32, device=cuda).contiguous()
    its = (_lib.ctypes.c_int * batch)(*iterations)
    lib = _lib.lib()
    nbytes = lib.csg_smooth_geometry_batch_workspace_bytes(n, batch)
    ws = _lib.workspace(nbytes, cuda)
    ref = torch.empty_like(geom)
    _lib.check(lib.csg_smooth_geometry_batch(
        _lib.ptr(geom), batch, _lib.ptr(topo.nbr_off), _lib.ptr(topo.nbr), n, its,
        _lib.ptr(ref), _lib.ptr(ws), nbytes, _lib.stream_handle(cuda)), "batch")
    pp = smoothing_patches(topo, geom[0], core=patches[0], rings=patches[1])
    out = torch.empty_like(geom)
    _lib.check(lib.csg_smooth_geometry_patched(
        _lib.ptr(geom), batch, n, its, pp.count, pp.rings, _lib.ptr(pp.perm), _lib.ptr(pp.vstart),
        _lib.ptr(pp.ring_end), _lib.ptr(pp.lverts), _lib.ptr(pp.aoff), _lib.ptr(pp.astart),
        _lib.ptr(pp.ladj), pp.lds_bytes, _lib.ptr(out), _lib.ptr(ws), nbytes,
        _lib.stream_handle(cuda)), "patched")
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
    off = np.ascontiguousarray(topo.nbr_off.cpu().numpy(), dtype=np.int32)
    nbr = np.ascontiguousarray(topo.nbr.cpu().numpy(), dtype=np.int32)
    host = geom.cpu().numpy()
    for b, k in enumerate(iterations):
        want = np.empty_like(host[b])
        O.lib().oracle_smooth_geometry(O._f(np.ascontiguousarray(host[b])), O._i(off),
                                       O._i(nbr), n, int(k), O._f(want))
        np.testing.assert_array_equal(ref[b].cpu().numpy(), want)


@pytest.mark.parametrize("level,counts,patches", [
    (6, (0, 1, 4, 7, 375, 166), (320, 4)),
    (5, (3, 9, 8, 13), (64, 7)),
    (3, (5, 2), (100000, 1)),
])
def test_value_smoothing_patched_equals_per_iteration(level, counts, patches, cuda):
    """Patch-local value smoothing (csg_smooth_values_patched) gives the bits of
    one launch per iteration (csg_smooth_values) for rows on two meshes with
    iteration counts of 0, below, at and between multiples of the rings per
    launch, in both distance forms."""

    import cat_surface_gpu as csg
    from cat_surface_gpu.stencil_builder import icosphere
    from cat_surface_gpu.surface_stencil import set_value_smoothing

    v, f = icosphere(level)
    n = v.shape[0]
    idx = np.repeat(np.arange(n, dtype=np.int32)[:, None], 3, axis=1)
    w = np.tile(np.array([1.0, 0.0, 0.0]), (n, 1))
    st = csg.SurfaceStencil(sphere_points=v, faces=f.astype(np.int32), surface_indices=idx,
                            surface_weights=w, sheet_indices=np.zeros((1, 1, 3), np.int32),
                            sheet_weights=np.zeros((1, 1, 3)), nx=1, ny=1,
                            source_points=n).to(cuda)
    rng = np.random.default_rng(level + 7)
    geom = torch.as_tensor(v[None] * 50.0 * (1.0 + 0.05 * rng.standard_normal((2, n, 1))),
                           dtype=torch.float32, device=cuda).contiguous()
    rows = torch.as_tensor(rng.standard_normal((len(counts), n)), dtype=torch.float64,
                           device=cuda)
    params = [(c, 1.0 + 0.3 * k) for k, c in enumerate(counts)]
    mesh_of = [k % 2 for k in range(len(counts))]
    for squared in (False, True):
        previous = set_value_smoothing(None)
        try:
            ref = st._smooth_values_batch(rows, geom, params, squared, mesh_of=mesh_of)
            set_value_smoothing(patches)
            out = st._smooth_values_batch(rows, geom, params, squared, mesh_of=mesh_of)
        finally:
            set_value_smoothing(previous)
        torch.cuda.synchronize()
        assert torch.equal(out, ref)


@pytest.mark.parametrize("level,grid", [(5, 32), (6, 32), (4, 64)])
def test_closest_triangles_pruned_equals_full_scan(level, grid, cuda):
    """Bounding-sphere pruning of the closest-triangle queries returns the faces,
    vertex indices and weights of the full candidate scan, including queries off
    the surface and exactly at vertices (ties)."""

    from cat_surface_gpu.locate import closest_triangles, set_closest_pruning
    from cat_surface_gpu.stencil_builder import icosphere, uv_points

    sv, sf = icosphere(level)
    rng = np.random.default_rng(level + grid)
    verts = torch.as_tensor(sv * (1.0 + 0.01 * rng.standard_normal((sv.shape[0], 1))),
                            device=cuda)
    faces = torch.as_tensor(sf, device=cuda)
    q = np.concatenate([uv_points(64, 32) * rng.uniform(0.97, 1.03, (2048, 1)),
                        sv[:500] * 0.99, sv[500:600]])
    q = torch.as_tensor(q, device=cuda)
    prev = set_closest_pruning(False)
    try:
        ref = closest_triangles(q, verts, faces, grid_size=grid, margin=1)
        set_closest_pruning(True)
        got = closest_triangles(q, verts, faces, grid_size=grid, margin=1)
    finally:
        set_closest_pruning(prev)
    for a, b in zip(got, ref):
        assert torch.equal(a, b)


@pytest.mark.parametrize("level,grid,spread", [
    (6, 64, 0.0), (6, 32, 0.0), (5, 64, 0.03), (6, 128, 0.0), (4, 16, 0.03), (6, 64, 0.025)])
def test_closest_triangles_block_equals_margin_search(level, grid, spread, cuda):
    """The block search on an index without margin (own cell, then the 3x3x3
    block where the own cell cannot decide) returns the faces, vertex indices
    and weights of the margin search, for map pixel directions on a perturbed
    unit mesh, points 2.5 % inside it (the surface stencil's queries, which
    mostly need the block), points at vertices (ties) and scattered points."""

    from cat_surface_gpu.locate import closest_triangles
    from cat_surface_gpu.stencil_builder import icosphere, uv_points

    sv, sf = icosphere(level)
    rng = np.random.default_rng(level * 7 + grid)
    pert = sv * (1.0 + 0.004 * rng.standard_normal((sv.shape[0], 1)))
    pert /= np.linalg.norm(pert, axis=1, keepdims=True)
    verts = torch.as_tensor(pert, device=cuda)
    faces = torch.as_tensor(sf, device=cuda)
    uv = uv_points(128, 64)
    scale = (1.0 if spread == 0.0 else rng.uniform(1.0 - spread, 1.0 + spread, (uv.shape[0], 1)))
    q = np.concatenate([uv * scale, sv[:400] * 0.975, pert[400:500]])
    q = torch.as_tensor(q, device=cuda)
    ref = closest_triangles(q, verts, faces, grid_size=grid, margin=1)
    got = closest_triangles(q, verts, faces, grid_size=grid, margin=1, block=True)
    for a, b in zip(got, ref):
        assert torch.equal(a, b)


@pytest.mark.parametrize("case", ["golden", "level6"])
def test_rotation_locator_skip_equals_full_scan(case, golden, cuda):
    """The locator's bounding-sphere skip leaves every location, cost, gradient
    term and refined rotation unchanged (bit for bit)."""

    import cat_surface_gpu as csg
    from cat_surface_gpu.stencil_builder import icosphere
    from cat_surface_gpu.surface_rotation import (axis_angle_matrix, refine_adjoint_launch,
                                                  rotation_matrix_np, set_locator_skip)

    rng = np.random.default_rng(21)
    if case == "golden":
        g = golden("rotation")
        idx = csg.RotationGridIndex.from_geometry(g["verts"], g["faces"], grid_size=32,
                                                  margin=1, device=cuda).to(cuda)
        inputs = idx.prepare_cost_inputs(g["src_pts"], g["src_vals"], g["tgt_vals"])
    else:
        v, f = icosphere(6)
        src, _ = icosphere(5)
        idx = csg.RotationGridIndex.from_geometry(v, f, grid_size=128, margin=1,
                                                  device=cuda).to(cuda)
        inputs = idx.prepare_cost_inputs(src.astype(np.float32), rng.normal(size=src.shape[0]),
                                         rng.normal(size=v.shape[0]))
    angles = torch.as_tensor(rng.uniform(-0.7, 0.7, size=(32, 3)))
    mats = np.stack([axis_angle_matrix(rng.normal(scale=0.3, size=3)) @
                     rotation_matrix_np(a) for a in angles.numpy()[:8]])
    out = {}
    prev = set_locator_skip(False)
    try:
        for skip in (False, True):
            set_locator_skip(skip)
            faces, bary = idx.locate(inputs, angles[:6])
            costs = idx.cost_batch_prepared(inputs, angles)
            terms = idx.cost_matrices(inputs, mats, gradient=True)
            refined = refine_adjoint_launch(idx, inputs,
                                            torch.as_tensor(mats.reshape(-1, 9), device=cuda),
                                            start_is_angles=False, steps=12)
            out[skip] = [faces, bary, costs, *terms, *refined[:3]]
    finally:
        set_locator_skip(prev)
    for a, b in zip(out[False], out[True]):
        assert torch.equal(a, b)
