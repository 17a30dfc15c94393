"""HIP DARTEL operators vs the CPU oracle and the reference golden vectors.

FP64 results must match the oracle to 1e-10 (the reference's own CUDA-vs-CPU
tolerance, tests/test_dartel_grid.py:312-341); FP32 results within the
tolerances stated per test.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL64 = dict(rtol=1e-10, atol=1e-10)


def _t(a, dev, dtype=torch.float64):
    return torch.as_tensor(np.asarray(a), dtype=dtype, device=dev)


def _np(t):
    return t.detach().double().cpu().numpy()


def _rand_map(nx, ny, scale, rng):
    return O.make_identity_map(nx, ny) + rng.standard_normal((2, ny, nx)) * scale


@pytest.fixture(scope="module")
def csg():
    import cat_surface_gpu

    return cat_surface_gpu


@pytest.mark.parametrize("shape", [(9, 7), (32, 24), (8, 1), (64, 48)])
def test_composition_and_jacobian(csg, cuda, shape):
    nx, ny = shape
    rng = np.random.default_rng(nx * 100 + ny)
    a = _rand_map(nx, ny, 1.7, rng)
    b = _rand_map(nx, ny, 0.8, rng)
    ja = 1.0 + 0.1 * rng.standard_normal((4, ny, nx))
    jb = 1.0 + 0.1 * rng.standard_normal((4, ny, nx))
    np.testing.assert_allclose(_np(csg.composition(_t(a, cuda), _t(b, cuda), device=cuda)),
                               O.composition(a, b), **TOL64)
    m, j = csg.composition_jacobian(_t(a, cuda), _t(ja, cuda), _t(b, cuda), _t(jb, cuda),
                                    device=cuda)
    om, oj = O.composition_jacobian(a, ja, b, jb)
    np.testing.assert_allclose(_np(m), om, **TOL64)
    np.testing.assert_allclose(_np(j), oj, **TOL64)


def test_grid_ops_match_golden(csg, cuda, golden):
    g = golden("grid_ops")
    for n in "abc":
        v = _t(g[f"{n}_v"], cuda)
        np.testing.assert_allclose(
            _np(csg.composition(_t(g[f"{n}_a"], cuda), _t(g[f"{n}_b"], cuda), device=cuda)),
            g[f"{n}_comp"], **TOL64)
        np.testing.assert_allclose(_np(csg.jacobian_of_displacement(v, 0.37, device=cuda)),
                                   g[f"{n}_jod"], **TOL64)
        m, j = csg.smalldef_jac(v, -0.25, device=cuda)
        np.testing.assert_allclose(_np(m), g[f"{n}_smalldef_map"], **TOL64)
        np.testing.assert_allclose(_np(j), g[f"{n}_smalldef_jac"], **TOL64)
        np.testing.assert_allclose(
            _np(csg.jac_div_smalldef(_t(g[f"{n}_ja"], cuda), v, 0.125, device=cuda)),
            g[f"{n}_jds"], **TOL64)
        for k in range(6):
            key = f"{n}_expdef_map_k{k}"
            if key not in g:
                continue
            m, j = csg.expdef(v * 0.5, k=k, return_jacobian=True, device=cuda)
            np.testing.assert_allclose(_np(m), g[key], **TOL64)
            np.testing.assert_allclose(_np(j), g[f"{n}_expdef_jac_k{k}"], **TOL64)
            np.testing.assert_allclose(_np(csg.expdef(v * 0.5, k=k, device=cuda)),
                                       g[f"{n}_expdef_only_k{k}"], **TOL64)
        for rtype in (0, 1, 2):
            np.testing.assert_allclose(
                _np(csg.regularization_operator(v, [1.3, 0.9, 0.125, 0.02, 0.001], rtype=rtype,
                                                device=cuda)),
                g[f"{n}_reg{rtype}"], **TOL64)


def test_objfun_and_squaring_match_golden(csg, cuda, golden):
    g = golden("objfun_squaring")
    for n in "ab":
        obj, rhs, hes = csg.initialise_objfun(
            _t(g[f"{n}_src"], cuda), _t(g[f"{n}_tgt"], cuda), _t(g[f"{n}_phi"], cuda),
            _t(g[f"{n}_jac"], cuda), device=cuda)
        np.testing.assert_allclose(float(obj), float(g[f"{n}_obj"]), rtol=1e-12)
        np.testing.assert_allclose(_np(rhs), g[f"{n}_rhs"], **TOL64)
        np.testing.assert_allclose(_np(hes), g[f"{n}_hes"], **TOL64)
        obj, rhs, hes = csg.initialise_objfun(
            _t(g[f"{n}_src"], cuda), _t(g[f"{n}_tgt"], cuda), _t(g[f"{n}_phi"], cuda),
            _t(g[f"{n}_jac"], cuda), _t(g[f"{n}_dist"], cuda), device=cuda)
        np.testing.assert_allclose(float(obj), float(g[f"{n}_obj_d"]), rtol=1e-12)
        np.testing.assert_allclose(_np(rhs), g[f"{n}_rhs_d"], **TOL64)
        for k in range(5):
            for save in (0, 1):
                out = csg.squaring_update(
                    _t(g[f"{n}_rhs"], cuda), _t(g[f"{n}_hes"], cuda), _t(g[f"{n}_sq_t0"], cuda),
                    _t(g[f"{n}_sq_j0"], cuda), k=k, save_transformation=bool(save),
                    return_scratch=True, device=cuda)
                tag = f"{n}_sq_k{k}_s{save}"
                for value, name in zip(out, ("b", "a", "map", "jac", "scratch")):
                    np.testing.assert_allclose(_np(value), g[f"{tag}_{name}"], **TOL64,
                                               err_msg=f"{tag}_{name}")


def test_solver_pieces_match_golden(csg, cuda, golden):
    g = golden("solver")
    p = [1.0, 1.0, 0.125, 0.0, 0.001]
    for n in ("even", "odd", "thin", "mid"):
        h, r, i = (_t(g[f"{n}_{x}"], cuda) for x in ("hes", "rhs", "init"))
        nx, ny = r.shape[2], r.shape[1]
        np.testing.assert_allclose(_np(csg.apply_membrane_system(h, r, p, device=cuda)),
                                   g[f"{n}_system"], **TOL64)
        np.testing.assert_allclose(_np(csg.relax_membrane(h, r, p, nit=4, device=cuda)),
                                   g[f"{n}_relax4"], **TOL64)
        np.testing.assert_allclose(
            _np(csg.relax_membrane(h, r, p, nit=2, initial=i, device=cuda)),
            g[f"{n}_relax2_init"], **TOL64)
        both = torch.cat([r, h])
        for tname, (tx, ty) in {"down": ((nx + 1) // 2, (ny + 1) // 2), "up": (2 * nx, 2 * ny),
                                "odd": (nx + 3, max(1, ny - 1))}.items():
            out = csg.resize_field(both, csg.GridSpec(nx=tx, ny=ty), device=cuda)
            np.testing.assert_allclose(_np(out), g[f"{n}_resize_{tname}"], **TOL64)
        if f"{n}_fmg_none" not in g:
            continue
        np.testing.assert_allclose(
            _np(csg.fmg2_membrane(h, r, p, cycles=2, nit=2, device=cuda)), g[f"{n}_fmg_none"],
            **TOL64)
        np.testing.assert_allclose(
            _np(csg.fmg2_membrane(h, r, p, cycles=3, nit=3, initial=i, device=cuda)),
            g[f"{n}_fmg_init"], **TOL64)
        z = torch.zeros_like(i)
        np.testing.assert_allclose(
            _np(csg.fmg2_membrane(h, r, p, cycles=3, nit=3, initial=z, device=cuda)),
            g[f"{n}_fmg_zero"], **TOL64)
        np.testing.assert_allclose(
            _np(csg.fmg2_membrane(h, r, p, cycles=3, nit=3, initial=z, device=cuda,
                                  assume_initial_nonzero=True)),
            g[f"{n}_fmg_zero_assume"], **TOL64)


@pytest.mark.parametrize("shape", [(9, 7), (5, 3), (3, 1), (1, 4), (7, 6)])
def test_odd_width_relaxation_matches_sequential_oracle(csg, cuda, shape):
    """Odd widths reproduce the reference's row-major in-place CPU sweep."""

    nx, ny = shape
    rng = np.random.default_rng(nx * 10 + ny)
    hes = rng.random((3, ny, nx))
    hes[0] += 2.0
    hes[1] += 2.0
    hes[2] *= 0.05
    rhs = rng.standard_normal((2, ny, nx))
    init = rng.standard_normal((2, ny, nx)) * 0.3
    p = [1.0, 1.0, 0.125, 0.0, 0.001]
    out = csg.relax_membrane(_t(hes, cuda), _t(rhs, cuda), p, nit=3, initial=_t(init, cuda),
                             device=cuda)
    np.testing.assert_allclose(_np(out), O.relax_membrane(hes, rhs, p, 3, init), **TOL64)
    fmg = csg.fmg2_membrane(_t(hes, cuda), _t(rhs, cuda), p, cycles=2, nit=2, device=cuda)
    np.testing.assert_allclose(_np(fmg), O.fmg2_membrane(hes, rhs, p, 2, 2), **TOL64)


def test_dartel_step_matches_golden(csg, cuda, golden):
    g = golden("dartel_step")
    p = [1.0, 1.0, 0.125, 0.02, 0.001]
    n = "a"
    src, tgt, v = (_t(g[f"{n}_{x}"], cuda) for x in ("src", "tgt", "v"))
    for code in (0, 1):
        for k in range(6):
            out, metrics = csg.dartel_step(src, tgt, v, k=k, params=p, lmreg=0.001, cycles=3,
                                           nit=3, code=code, device=cuda)
            np.testing.assert_allclose(_np(out), g[f"{n}_c{code}_k{k}_v"], **TOL64,
                                       err_msg=f"code {code} k {k}")
            np.testing.assert_allclose(_np(metrics), g[f"{n}_c{code}_k{k}_m"], rtol=1e-10)
    zero = torch.zeros_like(v)
    for assume, key in ((True, "zero_assume"), (False, "zero")):
        out, metrics = csg.dartel_step(src, tgt, zero, k=2, params=p, lmreg=0.001, code=1,
                                       device=cuda, assume_initial_nonzero=assume)
        np.testing.assert_allclose(_np(out), g[f"{n}_{key}_v"], **TOL64)


def test_dartel_step_with_distortion_matches_oracle(csg, cuda):
    rng = np.random.default_rng(5)
    nx, ny = 16, 12
    src, tgt = rng.random((ny, nx)), rng.random((ny, nx))
    v = rng.standard_normal((2, ny, nx)) * 0.2
    dist = 0.5 + rng.random((ny, nx))
    p = [1.0, 1.0, 0.125, 0.0, 0.0]
    out, metrics = csg.dartel_step(_t(src, cuda), _t(tgt, cuda), _t(v, cuda), k=3, params=p,
                                   distortion=_t(dist, cuda), device=cuda)
    ov, om = O.dartel_step(src, tgt, v, 3, p, 1e-3, 3, 3, 1, dist)
    np.testing.assert_allclose(_np(out), ov, **TOL64)
    np.testing.assert_allclose(_np(metrics), om, rtol=1e-10)


@pytest.mark.parametrize("graph", [False, True])
def test_solve_matches_golden(csg, cuda, golden, graph):
    g = golden("solve")
    for n in "ab":
        src, tgt = _t(g[f"{n}_src"], cuda), _t(g[f"{n}_tgt"], cuda)
        if graph:
            res = csg.solve_dartel_maps(src, tgt, device=cuda, cuda_graph=True)
            flow_key, met_key = f"{n}_flow_assume", f"{n}_metrics_assume"
        else:
            res = csg.solve_dartel_maps(src, tgt, device=cuda)
            flow_key, met_key = f"{n}_flow", f"{n}_metrics"
        np.testing.assert_allclose(_np(res.flow), g[flow_key], rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(_np(res.metrics), g[met_key], rtol=1e-9, atol=1e-12)
    csg.clear_dartel_cuda_graph_cache()


def test_graph_replay_is_stable(csg, cuda, golden):
    """Replays of a captured solve with eager work between them stay identical."""

    g = golden("solve")
    src, tgt = _t(g["a_src"], cuda), _t(g["a_tgt"], cuda)
    first = csg.solve_dartel_maps(src, tgt, device=cuda, cuda_graph=True).flow
    for _ in range(3):
        eager = csg.solve_dartel_maps(src, tgt, device=cuda, assume_initial_nonzero=True).flow
        torch.cuda.synchronize()
        replay = csg.solve_dartel_maps(src, tgt, device=cuda, cuda_graph=True).flow
        assert torch.equal(replay, first)
        torch.testing.assert_close(replay, eager, rtol=0, atol=0)
    csg.clear_dartel_cuda_graph_cache()


def test_fp32_solve_within_stage_tolerance(csg, cuda, golden):
    """FP32 solve vs FP64: flow displacement relative L2 within 1e-4."""

    g = golden("solve")
    src, tgt = g["b_src"], g["b_tgt"]
    r64 = csg.solve_dartel_maps(_t(src, cuda), _t(tgt, cuda), device=cuda)
    r32 = csg.solve_dartel_maps(_t(src, cuda, torch.float32), _t(tgt, cuda, torch.float32),
                                dtype=torch.float32, device=cuda)
    ident = O.make_identity_map(src.shape[2], src.shape[1])
    d64 = _np(r64.flow) - ident
    d32 = _np(r32.flow) - ident
    rel = np.linalg.norm(d32 - d64) / np.linalg.norm(d64)
    assert rel <= 1e-4, rel
