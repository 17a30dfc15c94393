"""The CPU oracle against the reference's golden vectors and closed forms.

Pins every DARTEL and warp function of oracle/csg_oracle.c: the vectors in
tests/golden/*.npz come from the reference's own CPU torch path (see
tests/golden/make_golden.py); the closed-form checks restate the reference's
in-test numpy oracles (tests/test_dartel_grid.py:36-107, :203-254).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as O

EXACT = dict(rtol=0.0, atol=0.0)
TIGHT = dict(rtol=1e-13, atol=1e-13)


def _bound_scalar(i, j, nx, ny):
    x = i % nx
    if ny == 1:
        return x, 0
    r = j % (2 * ny)
    return x, (2 * ny - r - 1 if r >= ny else r)


def test_mixed_boundary_matches_scalar_definition():
    nx, ny = 5, 4
    ix = np.array([-7, -1, 0, 4, 5, 12])
    iy = np.array([-9, -1, 0, 3, 4, 11])
    ox, oy = O.bound(ix, iy, nx, ny)
    assert list(zip(ox.tolist(), oy.tolist())) == [_bound_scalar(int(a), int(b), nx, ny)
                                                    for a, b in zip(ix, iy)]
    ox, oy = O.bound(ix, iy, 3, 1)
    assert oy.tolist() == [0] * 6


def _composition_scalar(a, b):
    """Scalar loop restatement of the composition (independent of the C code)."""

    ny, nx = a.shape[1:]
    out = np.empty_like(a)
    for y in range(ny):
        for x in range(nx):
            xx, yy = a[0, y, x] - 1.0, a[1, y, x] - 1.0
            ix, iy = math.floor(xx), math.floor(yy)
            dx1, dy1 = xx - ix, yy - iy
            for comp, period in ((0, nx), (1, ny)):
                vals = []
                for ox, oy in ((0, 0), (1, 0), (0, 1), (1, 1)):
                    cx, cy = _bound_scalar(ix + ox, iy + oy, nx, ny)
                    vals.append(b[comp, cy, cx] - 1.0)
                k22, k12, k21, k11 = vals
                k12 -= math.floor((k12 - k22) / period + 0.5) * period
                k21 -= math.floor((k21 - k22) / period + 0.5) * period
                k11 -= math.floor((k11 - k22) / period + 0.5) * period
                out[comp, y, x] = ((k22 * (1 - dx1) + k12 * dx1) * (1 - dy1)
                                   + (k21 * (1 - dx1) + k11 * dx1) * dy1 + 1.0)
    return out


def test_composition_matches_scalar_restatement():
    rng = np.random.default_rng(7)
    nx, ny = 7, 5
    a = O.make_identity_map(nx, ny) + rng.standard_normal((2, ny, nx)) * 1.7
    b = O.make_identity_map(nx, ny) + rng.standard_normal((2, ny, nx)) * 0.8
    np.testing.assert_allclose(O.composition(a, b), _composition_scalar(a, b), rtol=1e-12,
                               atol=1e-12)


def test_identity_is_neutral_for_composition():
    rng = np.random.default_rng(3)
    nx, ny = 12, 8
    ident = O.make_identity_map(nx, ny)
    b = ident + rng.standard_normal((2, ny, nx)) * 0.3
    np.testing.assert_allclose(O.composition(ident, b), b, **TIGHT)


def test_expdef_identity_limit_and_small_flow_jacobian():
    nx, ny = 6, 4
    m, j = O.expdef(np.zeros((2, ny, nx)), 3, True)
    np.testing.assert_allclose(m, O.make_identity_map(nx, ny), **TIGHT)
    expected = np.zeros((4, ny, nx))
    expected[0] = expected[3] = 1.0
    np.testing.assert_allclose(j, expected, **TIGHT)
    # central-difference Jacobian vs explicit differences (tests/test_dartel_grid.py:225-254)
    nx, ny = 8, 6
    x = np.arange(nx)[None, :]
    y = np.arange(ny)[:, None]
    v = np.stack([np.broadcast_to(0.03 * np.sin(2 * np.pi * x / nx), (ny, nx)),
                  np.broadcast_to(0.02 * np.cos(2 * np.pi * y / (2 * ny)), (ny, nx))]).copy()
    jac = O.jacobian_of_displacement(v)
    exp = np.empty((4, ny, nx))
    for r in range(ny):
        for c in range(nx):
            xm, ym = _bound_scalar(c - 1, r, nx, ny)
            xp, yp = _bound_scalar(c + 1, r, nx, ny)
            x0, y0 = _bound_scalar(c, r - 1, nx, ny)
            x1, y1 = _bound_scalar(c, r + 1, nx, ny)
            exp[0, r, c] = (v[0, yp, xp] - v[0, ym, xm]) / 2 + 1
            exp[1, r, c] = (v[1, yp, xp] - v[1, ym, xm]) / 2
            exp[2, r, c] = (v[0, y1, x1] - v[0, y0, x0]) / 2
            exp[3, r, c] = (v[1, y1, x1] - v[1, y0, x0]) / 2 + 1
    np.testing.assert_allclose(jac, exp, rtol=1e-12, atol=1e-12)


def test_expdef_inverse_pair_is_near_identity():
    """exp(v)∘exp(-v) ≈ id for a smooth small field (group property)."""

    nx, ny = 32, 16
    x = np.arange(nx)[None, :] / nx
    y = np.arange(ny)[:, None] / ny
    v = np.stack([np.broadcast_to(0.4 * np.sin(2 * np.pi * x) * np.cos(np.pi * y), (ny, nx)),
                  np.broadcast_to(0.3 * np.cos(2 * np.pi * x) * np.sin(np.pi * y), (ny, nx))])
    fwd = O.expdef(v, 6)
    inv = O.expdef(-v, 6)
    both = O.composition(inv, fwd)
    # rows next to the reflected y boundary see the mirrored field; check the interior
    err = np.abs(both - O.make_identity_map(nx, ny))[:, 3:-3]
    assert err.max() < 3e-3, err.max()


def test_membrane_system_is_symmetric_positive_definite():
    """(L + A) assembled column by column is SPD, so FMG solves a well-posed system."""

    rng = np.random.default_rng(11)
    nx, ny = 6, 4
    hes = rng.random((3, ny, nx))
    hes[0] += 2.0
    hes[1] += 2.0
    hes[2] *= 0.05
    p = [1.0, 1.0, 0.125, 0.0, 0.001]
    n = 2 * nx * ny
    mat = np.empty((n, n))
    for col in range(n):
        e = np.zeros(n)
        e[col] = 1.0
        mat[:, col] = O.apply_membrane_system(hes, e.reshape(2, ny, nx), p).reshape(-1)
    # x is circular and y reflected: the operator is symmetric
    np.testing.assert_allclose(mat, mat.T, rtol=1e-12, atol=1e-12)
    assert np.linalg.eigvalsh(mat).min() > 0


def test_relaxation_converges_to_dense_solution():
    rng = np.random.default_rng(12)
    nx, ny = 8, 4
    hes = rng.random((3, ny, nx))
    hes[0] += 2.0
    hes[1] += 2.0
    hes[2] *= 0.05
    rhs = rng.standard_normal((2, ny, nx))
    p = [1.0, 1.0, 0.125, 0.0, 0.001]
    n = 2 * nx * ny
    mat = np.empty((n, n))
    for col in range(n):
        e = np.zeros(n)
        e[col] = 1.0
        mat[:, col] = O.apply_membrane_system(hes, e.reshape(2, ny, nx), p).reshape(-1)
    exact = np.linalg.solve(mat, rhs.reshape(-1)).reshape(2, ny, nx)
    # det * 1.0000000001 perturbs the fixed point by ~1e-10 relative
    np.testing.assert_allclose(O.relax_membrane(hes, rhs, p, 200), exact, rtol=1e-8, atol=1e-9)
    fmg = O.fmg2_membrane(hes, rhs, p, cycles=40, nit=3)
    np.testing.assert_allclose(fmg, exact, rtol=1e-7, atol=1e-8)


def test_resize_weights_are_a_partition_of_unity():
    rng = np.random.default_rng(2)
    const = np.full((1, 9, 12), 3.25)
    for tnx, tny in ((6, 5), (24, 18), (13, 4), (1, 1)):
        np.testing.assert_allclose(O.resize_field(const, tnx, tny), 3.25, rtol=1e-14)
    assert O.resize_field(rng.random((2, 5, 6)), 3, 3).shape == (2, 3, 3)


def test_default_parameters_match_cat_mu_schedule():
    p = O.default_dartel_parameters(6)
    assert p[:, 2].tolist() == [0.125, 0.125, 0.125, 0.125, 0.1, 0.1]
    assert p[:, 0].tolist() == [1.0] * 6


# ----------------------------------------------------------- golden vectors


def test_grid_ops_golden(golden):
    g = golden("grid_ops")
    for n in "abc":
        np.testing.assert_allclose(O.composition(g[f"{n}_a"], g[f"{n}_b"]), g[f"{n}_comp"],
                                   **EXACT)
        np.testing.assert_allclose(g[f"{n}_comp_unopt"], g[f"{n}_comp"], **EXACT)
        m, j = O.composition_jacobian(g[f"{n}_a"], g[f"{n}_ja"], g[f"{n}_b"], g[f"{n}_jb"])
        np.testing.assert_allclose(m, g[f"{n}_cjac_map"], **EXACT)
        np.testing.assert_allclose(j, g[f"{n}_cjac_jac"], **EXACT)
        np.testing.assert_allclose(O.jacobian_of_displacement(g[f"{n}_v"], 0.37), g[f"{n}_jod"],
                                   **EXACT)
        m, j = O.smalldef_jac(g[f"{n}_v"], -0.25)
        np.testing.assert_allclose(m, g[f"{n}_smalldef_map"], **EXACT)
        np.testing.assert_allclose(j, g[f"{n}_smalldef_jac"], **EXACT)
        np.testing.assert_allclose(O.jac_div_smalldef(g[f"{n}_ja"], g[f"{n}_v"], 0.125),
                                   g[f"{n}_jds"], **EXACT)
        for k in range(6):
            if f"{n}_expdef_map_k{k}" not in g:
                continue
            m, j = O.expdef(g[f"{n}_v"] * 0.5, k, True)
            np.testing.assert_allclose(m, g[f"{n}_expdef_map_k{k}"], **EXACT)
            np.testing.assert_allclose(j, g[f"{n}_expdef_jac_k{k}"], **EXACT)
            np.testing.assert_allclose(O.expdef(g[f"{n}_v"] * 0.5, k), g[f"{n}_expdef_only_k{k}"],
                                       **EXACT)
        for r in range(3):
            np.testing.assert_allclose(
                O.regularization_operator(g[f"{n}_v"], [1.3, 0.9, 0.125, 0.02, 0.001], r),
                g[f"{n}_reg{r}"], **EXACT)


def test_objfun_and_squaring_golden(golden):
    g = golden("objfun_squaring")
    for n in "ab":
        o, r, h = O.initialise_objfun(g[f"{n}_src"], g[f"{n}_tgt"], g[f"{n}_phi"], g[f"{n}_jac"])
        np.testing.assert_allclose(o, g[f"{n}_obj"], rtol=1e-14)
        np.testing.assert_allclose(r, g[f"{n}_rhs"], **EXACT)
        np.testing.assert_allclose(h, g[f"{n}_hes"], **EXACT)
        o, r, h = O.initialise_objfun(g[f"{n}_src"], g[f"{n}_tgt"], g[f"{n}_phi"], g[f"{n}_jac"],
                                      g[f"{n}_dist"])
        np.testing.assert_allclose(o, g[f"{n}_obj_d"], rtol=1e-14)
        np.testing.assert_allclose(r, g[f"{n}_rhs_d"], **EXACT)
        np.testing.assert_allclose(h, g[f"{n}_hes_d"], **EXACT)
        for k in range(5):
            for s in (0, 1):
                out = O.squaring_update(g[f"{n}_rhs"], g[f"{n}_hes"], g[f"{n}_sq_t0"],
                                        g[f"{n}_sq_j0"], k, bool(s))
                tag = f"{n}_sq_k{k}_s{s}"
                for value, name in zip(out, ("b", "a", "map", "jac", "scratch")):
                    np.testing.assert_allclose(value, g[f"{tag}_{name}"], **EXACT,
                                               err_msg=f"{tag}_{name}")


def test_solver_golden(golden):
    g = golden("solver")
    p = [1.0, 1.0, 0.125, 0.0, 0.001]
    for n in ("even", "odd", "thin", "mid"):
        h, r, i = g[f"{n}_hes"], g[f"{n}_rhs"], g[f"{n}_init"]
        nx, ny = r.shape[2], r.shape[1]
        np.testing.assert_allclose(O.apply_membrane_system(h, r, p), g[f"{n}_system"], **EXACT)
        np.testing.assert_allclose(O.relax_membrane(h, r, p, 4), g[f"{n}_relax4"], **EXACT)
        np.testing.assert_allclose(O.relax_membrane(h, r, p, 2, i), g[f"{n}_relax2_init"],
                                   **EXACT)
        for t, (tx, ty) in {"down": ((nx + 1) // 2, (ny + 1) // 2), "up": (2 * nx, 2 * ny),
                            "odd": (nx + 3, max(1, ny - 1))}.items():
            np.testing.assert_allclose(O.resize_field(np.concatenate([r, h]), tx, ty),
                                       g[f"{n}_resize_{t}"], **EXACT)
        if f"{n}_fmg_none" in g:
            np.testing.assert_allclose(O.fmg2_membrane(h, r, p, 2, 2), g[f"{n}_fmg_none"],
                                       **EXACT)
            np.testing.assert_allclose(O.fmg2_membrane(h, r, p, 3, 3, i), g[f"{n}_fmg_init"],
                                       **EXACT)
            z = np.zeros_like(i)
            np.testing.assert_allclose(O.fmg2_membrane(h, r, p, 3, 3, z), g[f"{n}_fmg_zero"],
                                       **EXACT)
            np.testing.assert_allclose(O.fmg2_membrane(h, r, p, 3, 3, z, True),
                                       g[f"{n}_fmg_zero_assume"], **EXACT)


def test_dartel_step_golden(golden):
    g = golden("dartel_step")
    p = [1.0, 1.0, 0.125, 0.02, 0.001]
    for n in "ab":
        cyc, nit = (3, 3) if n == "a" else (1, 1)
        for c in (0, 1):
            for k in range(6):
                v, m = O.dartel_step(g[f"{n}_src"], g[f"{n}_tgt"], g[f"{n}_v"], k, p, 0.001, cyc,
                                     nit, c)
                np.testing.assert_allclose(v, g[f"{n}_c{c}_k{k}_v"], **EXACT)
                np.testing.assert_allclose(m, g[f"{n}_c{c}_k{k}_m"], rtol=1e-13)
        z = np.zeros_like(g[f"{n}_v"])
        for assume, key in ((True, "zero_assume"), (False, "zero")):
            v, m = O.dartel_step(g[f"{n}_src"], g[f"{n}_tgt"], z, 2, p, 0.001, 3, 3, 1, None,
                                 assume)
            np.testing.assert_allclose(v, g[f"{n}_{key}_v"], **EXACT)


@pytest.mark.parametrize("name", ["a", "b"])
def test_solve_golden(golden, name):
    g = golden("solve")
    np.testing.assert_allclose(O.default_dartel_parameters(6), g["params"], **EXACT)
    f, m = O.solve_dartel_maps(g[f"{name}_src"], g[f"{name}_tgt"])
    np.testing.assert_allclose(f, g[f"{name}_flow"], **EXACT)
    np.testing.assert_allclose(m, g[f"{name}_metrics"], rtol=1e-13)
    f, m = O.solve_dartel_maps(g[f"{name}_src"], g[f"{name}_tgt"], assume_initial_nonzero=True)
    np.testing.assert_allclose(f, g[f"{name}_flow_assume"], **EXACT)


def test_warp_golden(golden):
    g = golden("warp")
    unit = O.normalise_points_f32(g["points"])
    np.testing.assert_allclose(O.apply_flow_to_sphere(unit, g["flow"], True), g["warp_inverse"],
                               rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(O.apply_flow_to_sphere(unit, g["flow"], False),
                               g["warp_forward"], rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(O.apply_flow_to_sphere(g["points"], g["flow"], True),
                               g["warp_cached"], rtol=1e-13, atol=1e-14)
    np.testing.assert_array_equal(O.wrap_displacement(g["wrap_in"]), g["wrap_out"])
    np.testing.assert_array_equal(O.average_xz_surfaces(g["avg_x"], g["avg_z"]), g["avg_out"])
    mats = np.stack([O.rotation_matrix_from_angles(a) for a in g["rot_angles"]])
    np.testing.assert_array_equal(mats, g["rot_mats"])
    np.testing.assert_array_equal(O.normalise_sphere(g["norm_in"]), g["norm_out"])
    np.testing.assert_array_equal(O.rotate_vertices(g["norm_out"], g["rot_angles"][1]),
                                  g["rotv_out"])


def test_warp_identity_flow_keeps_points():
    """A zero displacement (identity flow) moves no point (closed form)."""

    rng = np.random.default_rng(4)
    pts = rng.standard_normal((200, 3)).astype(np.float32)
    unit = O.normalise_points_f32(pts)
    flow = O.make_identity_map(32, 16)
    out = O.apply_flow_to_sphere(unit, flow, True)
    np.testing.assert_allclose(out, unit.astype(np.float64), atol=2e-6)
    np.testing.assert_allclose(np.linalg.norm(out, axis=1), 1.0, rtol=1e-12)
