"""Pin the surface-side oracle (oracle/csg_oracle_surface.c) on the CPU.

Each oracle function is checked against the reference's golden vectors
(tests/golden/surface.npz, rotation.npz: written from the reference's CPU
path by tests/golden/make_golden.py) and, where they leave something open,
against closed forms: partitions of unity, the identity rotation, exact
barycentric reconstruction and brute-force minimality.  Float32 geometry is
compared within float32 rounding (the reference sums with torch CPU
reductions of unspecified order); float64 value paths within 1e-12.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O


@pytest.fixture(scope="module")
def surf(golden):
    return golden("surface")


@pytest.fixture(scope="module")
def rot(golden):
    return golden("rotation")


def test_resample_is_bit_exact(surf):
    out = O.resample_vertices_f32(surf["surface"], surf["surf_idx"], surf["surf_w"])
    np.testing.assert_array_equal(out, surf["resampled"])


def test_edge_mean_and_heat_parameters(surf):
    assert abs(O.edge_mean(surf["surface"], surf["faces"]) - float(surf["edge_mean"])) < 1e-12
    it, sigma = O.heat_parameters(surf["surface"], surf["faces"], 50.0)
    assert it == int(surf["heat50_it"])
    assert abs(sigma - float(surf["heat50_sigma"])) < 1e-12
    # closed form: iterations = ceil(fwhm^2 * 0.541011 / e^2), sigma = fwhm * 0.7355345 / sqrt(it)
    e = O.edge_mean(surf["surface"], surf["faces"])
    assert it == int(np.ceil(50.0 ** 2 * 0.541011 / e ** 2))
    assert abs(sigma - 50.0 * 0.7355345 / np.sqrt(it)) < 1e-12


@pytest.mark.parametrize("fwhm,key", [(50.0, "smooth_geom50"), (15.0, "smooth_geom15")])
def test_smooth_geometry(surf, fwhm, key):
    out = O.smooth_geometry(surf["surface"], surf["faces"], fwhm)
    np.testing.assert_allclose(out, surf[key], rtol=2e-7, atol=1e-5)


def test_smooth_values(surf):
    out = O.smooth_values(surf["values"], surf["surface"], surf["faces"], 5.0)
    np.testing.assert_allclose(out, surf["smooth_vals5"], rtol=1e-12, atol=1e-14)
    many0 = O.smooth_values(surf["values"], surf["surface"], surf["faces"], 5.0)
    many1 = O.smooth_values(surf["values"], surf["surface"], surf["faces"], 5.0 / 3.0)
    np.testing.assert_allclose(many0, surf["smooth_many0"], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(many1, surf["smooth_many1"], rtol=1e-12, atol=1e-14)
    # a constant field is a fixed point of the normalised heat kernel
    const = O.smooth_values(np.full(surf["values"].shape, 3.25), surf["surface"], surf["faces"],
                            5.0)
    np.testing.assert_allclose(const, 3.25, rtol=1e-14)


def test_vertex_normals(surf):
    n = O.vertex_normals(surf["surface"], surf["faces"])
    np.testing.assert_allclose(n, surf["normals"], atol=2e-6)
    np.testing.assert_allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-6)


def test_sheet_gather_and_type5_maps(surf):
    sheet = O.sheet_gather(surf["values"], surf["sheet_idx"].reshape(-1, 3),
                           surf["sheet_w"].reshape(-1, 3)).reshape(16, 32)
    np.testing.assert_allclose(sheet, surf["sheet_of_values"], rtol=1e-15, atol=0)
    maps = O.curvature_type5_to_sheet_many(surf["surface"], surf["faces"], surf["sheet_idx"],
                                           surf["sheet_w"], (5.0 / 9.0, 5.0 / 27.0))
    np.testing.assert_allclose(maps[0], surf["type5_sheet0"], atol=2e-6)
    np.testing.assert_allclose(maps[1], surf["type5_sheet1"], atol=2e-6)


def test_minmax_normalise_closed_form():
    x = np.array([3.0, -1.0, 5.0, 2.0])
    np.testing.assert_allclose(O.minmax_normalise(x), (x + 1.0) / 6.0, rtol=1e-15)


def test_grid_index_matches_reference(rot):
    off, faces = O.grid_index(rot["verts"], rot["faces"], grid_size=32, margin=1)
    np.testing.assert_array_equal(off, rot["cand_offsets"])
    np.testing.assert_array_equal(faces, rot["cand_faces"])


def test_seed_grid_and_rotation_costs(rot):
    np.testing.assert_array_equal(O.official_seed_grid_angles(), rot["seeds"])
    costs, lf, lb = O.rotation_costs(rot["src_pts"], rot["src_vals"], rot["tgt_vals"],
                                     rot["verts"], rot["faces"], rot["seeds"], 32, 1,
                                     return_location=True)
    # float32 location, float64 sums in another order than torch's: 5.7e-8 relative
    np.testing.assert_allclose(costs, rot["seed_costs"], rtol=2e-7)
    assert int(np.argmin(costs)) == int(np.argmin(rot["seed_costs"]))
    np.testing.assert_array_equal(lf[:5, :200], rot["loc_faces"])
    np.testing.assert_allclose(lb[:5, :200], rot["loc_bary"], rtol=0, atol=2e-6)


def test_identity_rotation_of_the_target_itself_costs_nothing():
    """Vertex values sampled at the vertices themselves cost 0 (reference test)."""

    from cat_surface_gpu.stencil_builder import icosphere

    v, f = icosphere(2)
    vals = np.arange(v.shape[0], dtype=np.float64)
    cost = O.rotation_costs(v.astype(np.float32), vals, vals, v, f, np.zeros((1, 3)), 16, 2)
    assert cost[0] < 1e-6


def test_nelder_mead_reaches_reference_minimum(rot):
    index = O.grid_index(rot["verts"], rot["faces"], grid_size=32, margin=1)

    def cost(angles):
        return O.rotation_costs(rot["src_pts"], rot["src_vals"], rot["tgt_vals"], rot["verts"],
                                rot["faces"], angles, 32, 1, index=index)

    angle, c, iterations = O.nelder_mead(cost, rot["seed_angle"], max_iter=60)
    np.testing.assert_allclose(angle, rot["nm_angle"], rtol=0, atol=1e-12)
    assert iterations == int(rot["nm_iters"])
    assert abs(c - float(rot["nm_cost"])) <= 2e-6 * abs(float(rot["nm_cost"]))


def test_rotation_matrix_layout(golden):
    w = golden("warp")
    for a, m in zip(w["rot_angles"], w["rot_mats"]):
        np.testing.assert_allclose(O.rotation_matrix_f32(a), m, rtol=1e-6, atol=1e-7)


def test_closest_triangles_reconstruct_and_minimise():
    """Barycentric weights reproduce the closest point; no other face is closer."""

    from cat_surface_gpu.stencil_builder import icosphere

    v, f = icosphere(2)
    rng = np.random.default_rng(5)
    q = rng.normal(size=(200, 3))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    q *= rng.uniform(0.9, 1.1, size=(200, 1))
    face, w = O.closest_triangles(v, f, q)
    np.testing.assert_allclose(w.sum(axis=1), 1.0, atol=1e-12)
    assert (w >= -1e-12).all()
    p = np.einsum("qc,qcd->qd", w, v[f[face]])
    d = np.linalg.norm(p - q, axis=1)
    # brute force over a dense barycentric sampling of every face
    s = np.linspace(0.0, 1.0, 9)
    a, b = np.meshgrid(s, s)
    keep = a + b <= 1.0
    bary = np.stack([1 - a[keep] - b[keep], a[keep], b[keep]], axis=1)
    samples = np.einsum("sc,fcd->fsd", bary, v[f]).reshape(-1, 3)
    dmin = np.min(np.linalg.norm(samples[None] - q[:, None], axis=2), axis=1)
    assert (d <= dmin + 1e-12).all()


def test_type2_curvature(surf):
    """Type-2 values and sheet (surface_stencil.py:552-644).  With the golden
    float32 normals the shape-operator fit reproduces the reference to float64
    rounding (4.4e-16); the oracle's own normals differ from torch's by one
    float32 ulp, which moves the normalised values by at most 1e-7."""

    import ctypes

    s, f = surf["surface"], surf["faces"]
    off, nbr = O.neighbour_csr(f, s.shape[0])
    raw = np.empty(s.shape[0])
    O.lib().oracle_curvature_type2(O._f(np.ascontiguousarray(s)),
                                   O._f(np.ascontiguousarray(surf["normals"])), O._i(off),
                                   O._i(nbr), ctypes.c_int(s.shape[0]), O._d(raw))
    exact = O.minmax_normalise(O.smooth_values(raw, s, f, 5.0))
    np.testing.assert_allclose(exact, surf["type2_values"], rtol=0, atol=2e-15)
    own = O.minmax_normalise(O.smooth_values(O.curvature_type2(s, f), s, f, 5.0))
    np.testing.assert_allclose(own, surf["type2_values"], rtol=0, atol=5e-7)
    sheet = O.curvature_to_sheet(s, f, surf["sheet_idx"], surf["sheet_w"], 5.0 / 81.0, 2)
    np.testing.assert_allclose(sheet, surf["type2_sheet"], rtol=0, atol=5e-7)


@pytest.mark.parametrize("subdivisions", [3, 4])
def test_type2_curvature_of_a_sphere_is_its_inverse_radius(subdivisions):
    """Closed form: both principal curvatures of a sphere of radius R are 1/R, so
    sqrt((k1^2 + k2^2) / 2) = 1/R up to the discretisation of the fit (2.1% at
    the worst vertex, 0.1% at the median)."""

    from oracle import pipeline as P

    v, f = P.icosphere(subdivisions)
    raw = O.curvature_type2((2.0 * v).astype(np.float32), f)
    np.testing.assert_allclose(raw, 0.5, rtol=0.025)
    assert abs(np.median(raw) - 0.5) < 1e-3


def test_single_fwhm_type5_sheet(surf):
    """curvature_to_sheet(curvtype=5) smooths with |d|^2 weights (smooth_values),
    the batched form with sums of squares; both give the reference's sheet."""

    sheet = O.curvature_to_sheet(surf["surface"], surf["faces"], surf["sheet_idx"],
                                 surf["sheet_w"], 5.0 / 9.0, 5)
    np.testing.assert_allclose(sheet, surf["type5_sheet_single"], rtol=0, atol=2e-6)
    np.testing.assert_allclose(sheet, surf["type5_sheet0"], rtol=0, atol=2e-6)
