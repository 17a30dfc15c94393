"""Pins of the oracle chain's own stages (oracle/pipeline.py, csg_oracle_pipeline.c).

The depth potential, colouring and rotation feature are compared with the
reference's fixture (tests/golden/rotation_feature.npz, written from
src/cat_surface_gpu/rotation_feature.py); the grid-accelerated closest-triangle
search with the brute-force search; the vectorised grid index with the
reference's fixture; the whole chain with its invariants on a small pair.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from oracle import pipeline as P


def test_greedy_colours_match_reference(golden):
    g = golden("rotation_feature")
    np.testing.assert_array_equal(P.greedy_colours(g["faces"], g["verts"].shape[0]),
                                  g["colours"])


def test_sor_depth_potential_matches_reference(golden):
    g = golden("rotation_feature")
    d = P.depth_potential(g["smoothed"], g["faces"])
    np.testing.assert_allclose(d.normals, g["normals"], rtol=0, atol=1e-14)
    np.testing.assert_allclose(d.areas, g["areas"], rtol=1e-14)
    scale = np.abs(g["depth"]).max()
    np.testing.assert_allclose(d.values, g["depth"], rtol=0, atol=1e-13 * scale)
    assert d.iterations == 1000
    assert abs(d.relative_residual - float(g["residual"])) <= 1e-14


def test_rotation_feature_matches_reference(golden):
    g = golden("rotation_feature")
    st = P.Stencil(sphere64=g["verts"], faces=g["faces"], surf_idx=g["surf_idx"],
                   surf_w=g["surf_w"], sheet_idx=None, sheet_w=None, unit_points=None)
    np.testing.assert_allclose(P.rotation_feature(st, g["surface"], 15.0), g["feature"],
                               rtol=0, atol=1e-13)


def test_cg_depth_potential_matches_reference(golden):
    """The reference's CG (rotation_feature.py:193-250, tolerance 1e-9): same
    iteration count, the solution within CG's rounding sensitivity at the stopping
    iteration (dot products in another order: 4.7e-11 relative), and the feature
    built on it."""

    g = golden("rotation_feature")
    cg = P.depth_potential(g["smoothed"], g["faces"], solver="pcg", tolerance=1e-9)
    assert cg.iterations == int(g["cg_iterations"]) == 119
    assert cg.relative_residual <= 1e-9
    assert abs(cg.relative_residual - float(g["cg_residual"])) <= 0.1 * float(g["cg_residual"])
    scale = np.abs(g["depth_cg"]).max()
    np.testing.assert_allclose(cg.values, g["depth_cg"], rtol=0, atol=1e-9 * scale)
    st = P.Stencil(sphere64=g["verts"], faces=g["faces"], surf_idx=g["surf_idx"],
                   surf_w=g["surf_w"], sheet_idx=None, sheet_w=None, unit_points=None)
    np.testing.assert_allclose(P.rotation_feature(st, g["surface"], 15.0, solver="pcg"),
                               g["feature_cg"], rtol=0, atol=1e-10)
    # both solvers solve the same system: SOR's converged values
    np.testing.assert_allclose(cg.values, g["depth"], rtol=0, atol=1e-8 * scale)


@pytest.mark.parametrize("seed", [0, 1])
def test_grid_closest_triangles_equal_brute_force(seed):
    rng = np.random.default_rng(seed)
    v, f = P.icosphere(3)
    v = v + rng.normal(scale=0.01, size=v.shape)
    q = rng.normal(size=(2000, 3))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    q *= rng.uniform(0.5, 1.5, size=(2000, 1))
    # queries on vertices and edge midpoints tie between faces: first face wins
    q = np.concatenate([q, v[:50], 0.5 * (v[f[:50, 0]] + v[f[:50, 1]])])
    fa, wa = P.closest_triangles(v, f, q)
    fb, wb = O.closest_triangles(v, f, q)
    np.testing.assert_array_equal(fa, fb)
    np.testing.assert_array_equal(wa, wb)


def test_vectorised_grid_index_matches_reference(golden):
    r = golden("rotation")
    off, cf = P.grid_index(r["verts"], r["faces"], 32, 1)
    np.testing.assert_array_equal(off, r["cand_offsets"])
    np.testing.assert_array_equal(cf, r["cand_faces"])


def test_icosphere_and_pixel_directions():
    from cat_surface_gpu.stencil_builder import icosphere, uv_points

    for k in range(5):
        v, f = P.icosphere(k)
        assert v.shape[0] == 10 * 4 ** k + 2 and f.shape[0] == 20 * 4 ** k
        np.testing.assert_allclose(np.linalg.norm(v, axis=1), 1.0, atol=1e-15)
        # a closed genus-0 surface: V - E + F = 2
        edges = np.unique(np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]),
                                  axis=1), axis=0)
        assert v.shape[0] - edges.shape[0] + f.shape[0] == 2
        gv, gf = icosphere(k)
        np.testing.assert_array_equal(v, gv)
        np.testing.assert_array_equal(f, gf)
    np.testing.assert_array_equal(P.uv_points(32, 16), uv_points(32, 16))


def test_neighbour_csr_is_sorted_unique_adjacency():
    v, f = P.icosphere(2)
    off, nbr = O.neighbour_csr(f, v.shape[0])
    for i in range(v.shape[0]):
        row = nbr[off[i]:off[i + 1]]
        assert np.all(np.diff(row) > 0)
        expect = np.unique(np.concatenate([f[np.any(f == i, axis=1)].reshape(-1)]))
        np.testing.assert_array_equal(row, expect[expect != i])


def test_stencil_reconstructs_points_and_map():
    """Resampled points are barycentric combinations of their located triangle;
    the map stencil's weights are a partition of unity on the unit sphere."""

    from cat_surface_gpu import synthetic

    p = synthetic.make_pair("icosphere", subdivisions=3, lmax=4, noise=0.0, seed=5)
    st = P.build_stencil(p.src_sphere, p.src_faces, normalise=True, nx=32, ny=16,
                         n_triangles=1280)
    assert np.all(st.surf_w >= -1e-12) and np.all(st.sheet_w >= -1e-12)
    np.testing.assert_allclose(st.surf_w.sum(1), 1.0, atol=1e-12)
    np.testing.assert_allclose(st.sheet_w.sum(-1), 1.0, atol=1e-12)
    np.testing.assert_allclose(np.linalg.norm(st.unit_points, axis=1), 1.0, atol=1e-6)


def test_chain_runs_end_to_end_on_a_small_pair():
    from cat_surface_gpu import synthetic

    p = synthetic.make_pair("icosphere", subdivisions=3, lmax=4, noise=0.0,
                            max_rotation_deg=10.0, seed=7)
    res = P.surfwarp(p.src_surface, p.src_sphere, p.src_faces, p.tgt_surface, p.tgt_sphere,
                     p.tgt_faces, steps=2, runs=2, avg=True, map_size=(64, 32), loop=2,
                     max_iter=40)
    assert res.vertices.shape == p.src_sphere.shape and res.vertices.dtype == np.float32
    np.testing.assert_allclose(np.linalg.norm(res.vertices, axis=1), 1.0, atol=1e-5)
    assert len(res.flows) == 3
    # the seed scan's first minimum and a refined cost no larger than it
    assert res.rotation.cost <= res.rotation.seed_costs[res.rotation.seed_index] + 1e-12
