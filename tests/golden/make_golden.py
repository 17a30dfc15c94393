#!/usr/bin/env python3
"""Write golden input/output vectors for the CAT_SurfWarp hot path.

This script is the only place that imports the reference implementation
(``ayakacxy/cat-surface-gpu``, mounted read-only at /root/reference).  It runs
the reference's CPU torch path (``device="cpu"``, ``kernel="torch"``) on small
seeded inputs and stores inputs and outputs as compressed ``.npz`` fixtures
beside this file.  The fixtures are data only; the reference itself never
travels with the repository.

Functions written here are those that no closed form, invariant or brute-force
check pins on its own (``initialise_objfun``, ``squaring_update`` and its
scratch return, ``jac_div_smalldef``, ``dartel_step``, ``fmg2_membrane``,
``solve_dartel_maps``, ``apply_flow_to_sphere``, ``_average_xz_surfaces``, the
heat-kernel smoothers, type-5 curvature sheets, rotation costs and the
Nelder-Mead trajectory).  Closed-form operators are written too, so the
oracle's restatement of every function is compared with the same vectors.

Run from the repository root:  python tests/golden/make_golden.py
"""

from __future__ import annotations

import math
import sys
from pathlib import Path

import numpy as np
import torch

REFERENCE_SRC = Path("/root/reference/src")
OUT = Path(__file__).resolve().parent

sys.path.insert(0, str(REFERENCE_SRC))

from cat_surface_gpu import dartel_grid as dg  # noqa: E402
from cat_surface_gpu import surface_warp as sw  # noqa: E402
from cat_surface_gpu import surface_stencil as ss  # noqa: E402
from cat_surface_gpu import surface_rotation as sr  # noqa: E402
from cat_surface_gpu import gpu_pipeline as gp  # noqa: E402

torch.set_num_threads(8)
F64 = torch.float64


def _save(name: str, **arrays) -> None:
    clean = {}
    for key, value in arrays.items():
        if isinstance(value, torch.Tensor):
            value = value.detach().cpu().numpy()
        clean[key] = np.asarray(value)
    np.savez_compressed(OUT / f"{name}.npz", **clean)
    size = (OUT / f"{name}.npz").stat().st_size
    print(f"wrote {name}.npz ({size} bytes, {len(clean)} arrays)", flush=True)


def _icosphere(subdivisions: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit icosphere by midpoint subdivision (numpy, test-fixture helper)."""

    t = (1.0 + 5.0**0.5) / 2.0
    verts = np.array(
        [
            (-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0),
            (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
            (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1),
        ],
        dtype=np.float64,
    )
    faces = np.array(
        [
            (0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
            (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
            (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
            (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1),
        ],
        dtype=np.int64,
    )
    verts /= np.linalg.norm(verts, axis=1, keepdims=True)
    for _ in range(subdivisions):
        edges = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]])
        key = np.sort(edges, axis=1)
        uniq, inverse = np.unique(key, axis=0, return_inverse=True)
        mids = verts[uniq[:, 0]] + verts[uniq[:, 1]]
        mids /= np.linalg.norm(mids, axis=1, keepdims=True)
        base = verts.shape[0]
        verts = np.concatenate([verts, mids])
        nf = faces.shape[0]
        m01 = base + inverse[:nf]
        m12 = base + inverse[nf : 2 * nf]
        m20 = base + inverse[2 * nf :]
        faces = np.concatenate(
            [
                np.stack([faces[:, 0], m01, m20], 1),
                np.stack([faces[:, 1], m12, m01], 1),
                np.stack([faces[:, 2], m20, m12], 1),
                np.stack([m01, m12, m20], 1),
            ]
        )
    return verts, faces.astype(np.int32)


def _random_map(spec, scale, generator):
    ident = dg.make_identity_map(spec, device="cpu")
    return ident + torch.randn(ident.shape, dtype=F64, generator=generator) * scale


def golden_grid_ops() -> None:
    g = torch.Generator().manual_seed(20261203)
    out = {}
    for name, (nx, ny) in {"a": (9, 7), "b": (32, 24), "c": (8, 1)}.items():
        spec = dg.GridSpec(nx=nx, ny=ny)
        a = _random_map(spec, 1.7, g)
        b = _random_map(spec, 0.8, g)
        ja = 1.0 + 0.1 * torch.randn((4, ny, nx), dtype=F64, generator=g)
        jb = 1.0 + 0.1 * torch.randn((4, ny, nx), dtype=F64, generator=g)
        comp = dg.composition(a, b, device="cpu")
        comp_u = dg.composition(a, b, device="cpu", optimized=False)
        cmap, cjac = dg.composition_jacobian(a, ja, b, jb, device="cpu")
        v = torch.randn((2, ny, nx), dtype=F64, generator=g) * 0.6
        jod = dg.jacobian_of_displacement(v, scale=0.37, device="cpu")
        sm_t, sm_j = dg.smalldef_jac(v, -0.25, device="cpu")
        jds = dg.jac_div_smalldef(ja, v, 0.125, device="cpu")
        out.update(
            {
                f"{name}_a": a, f"{name}_b": b, f"{name}_ja": ja, f"{name}_jb": jb,
                f"{name}_comp": comp, f"{name}_comp_unopt": comp_u,
                f"{name}_cjac_map": cmap, f"{name}_cjac_jac": cjac,
                f"{name}_v": v, f"{name}_jod": jod,
                f"{name}_smalldef_map": sm_t, f"{name}_smalldef_jac": sm_j,
                f"{name}_jds": jds,
            }
        )
        for k in (range(0, 6) if name != "b" else (0, 2, 5)):
            phi, jac = dg.expdef(v * 0.5, k=k, return_jacobian=True, device="cpu")
            phi_only = dg.expdef(v * 0.5, k=k, device="cpu")
            out[f"{name}_expdef_map_k{k}"] = phi
            out[f"{name}_expdef_jac_k{k}"] = jac
            out[f"{name}_expdef_only_k{k}"] = phi_only
        params = [1.3, 0.9, 0.125, 0.02, 0.001]
        for rtype in (0, 1, 2):
            out[f"{name}_reg{rtype}"] = dg.regularization_operator(
                v, params, rtype=rtype, device="cpu"
            )
    _save("grid_ops", **out)


def golden_objfun_squaring() -> None:
    g = torch.Generator().manual_seed(7301)
    out = {}
    for name, (nx, ny) in {"a": (9, 7), "b": (32, 24)}.items():
        spec = dg.GridSpec(nx=nx, ny=ny)
        src = torch.rand((ny, nx), dtype=F64, generator=g)
        tgt = torch.rand((ny, nx), dtype=F64, generator=g)
        v = torch.randn((2, ny, nx), dtype=F64, generator=g) * 0.8
        dist = 0.5 + torch.rand((ny, nx), dtype=F64, generator=g)
        phi, jac = dg.expdef(v, k=2, return_jacobian=True, device="cpu")
        obj, rhs, hes = dg.initialise_objfun(src, tgt, phi, jac, device="cpu")
        obj_d, rhs_d, hes_d = dg.initialise_objfun(src, tgt, phi, jac, dist, device="cpu")
        out.update(
            {
                f"{name}_src": src, f"{name}_tgt": tgt, f"{name}_v": v,
                f"{name}_dist": dist, f"{name}_phi": phi, f"{name}_jac": jac,
                f"{name}_obj": obj, f"{name}_rhs": rhs, f"{name}_hes": hes,
                f"{name}_obj_d": obj_d, f"{name}_rhs_d": rhs_d, f"{name}_hes_d": hes_d,
            }
        )
        t0, j0 = dg.smalldef_jac(v, -0.25, device="cpu")
        for k in range(0, 5):
            for save in (False, True):
                b2, a2, m2, jj2, scratch = dg.squaring_update(
                    rhs, hes, t0, j0, k=k, save_transformation=save,
                    return_scratch=True, device="cpu", kernel="torch",
                )
                tag = f"{name}_sq_k{k}_s{int(save)}"
                out[f"{tag}_b"] = b2
                out[f"{tag}_a"] = a2
                out[f"{tag}_map"] = m2
                out[f"{tag}_jac"] = jj2
                out[f"{tag}_scratch"] = scratch
        out[f"{name}_sq_t0"] = t0
        out[f"{name}_sq_j0"] = j0
    _save("objfun_squaring", **out)


def golden_solver() -> None:
    g = torch.Generator().manual_seed(991)
    out = {}
    cases = {"even": (32, 24), "odd": (9, 7), "thin": (6, 1), "mid": (64, 32)}
    for name, (nx, ny) in cases.items():
        spec = dg.GridSpec(nx=nx, ny=ny)
        hes = torch.rand((3, ny, nx), dtype=F64, generator=g)
        hes[0] += 2.0
        hes[1] += 2.0
        hes[2] *= 0.05
        rhs = torch.randn((2, ny, nx), dtype=F64, generator=g)
        init = torch.randn((2, ny, nx), dtype=F64, generator=g) * 0.3
        params = [1.0, 1.0, 0.125, 0.0, 0.001]
        out[f"{name}_hes"] = hes
        out[f"{name}_rhs"] = rhs
        out[f"{name}_init"] = init
        out[f"{name}_system"] = dg.apply_membrane_system(
            hes, rhs, params, kernel="torch", device="cpu"
        )
        out[f"{name}_relax4"] = dg.relax_membrane(
            hes, rhs, params, nit=4, kernel="torch", device="cpu"
        )
        out[f"{name}_relax2_init"] = dg.relax_membrane(
            hes, rhs, params, nit=2, initial=init, kernel="torch", device="cpu"
        )
        for tname, (tx, ty) in {"down": ((nx + 1) // 2, (ny + 1) // 2), "up": (2 * nx, 2 * ny), "odd": (nx + 3, max(1, ny - 1))}.items():
            out[f"{name}_resize_{tname}"] = dg.resize_field(
                torch.cat([rhs, hes]), dg.GridSpec(nx=tx, ny=ty), kernel="torch", device="cpu"
            )
        if nx % 2 == 0 or nx == 1:
            out[f"{name}_fmg_none"] = dg.fmg2_membrane(
                hes, rhs, params, cycles=2, nit=2, kernel="torch", device="cpu"
            )
            out[f"{name}_fmg_init"] = dg.fmg2_membrane(
                hes, rhs, params, cycles=3, nit=3, initial=init, kernel="torch", device="cpu"
            )
            out[f"{name}_fmg_zero"] = dg.fmg2_membrane(
                hes, rhs, params, cycles=3, nit=3, initial=torch.zeros_like(init),
                kernel="torch", device="cpu",
            )
            out[f"{name}_fmg_zero_assume"] = dg.fmg2_membrane(
                hes, rhs, params, cycles=3, nit=3, initial=torch.zeros_like(init),
                kernel="torch", device="cpu", assume_initial_nonzero=True,
            )
    _save("solver", **out)


def _smooth_maps(nx, ny, generator, n_cos=24):
    y = (torch.arange(ny, dtype=F64) + 0.5).view(ny, 1) / ny
    x = (torch.arange(nx, dtype=F64) + 0.5).view(1, nx) / nx
    value = torch.zeros((ny, nx), dtype=F64)
    for _ in range(n_cos):
        fx = float(torch.randint(0, 4, (1,), generator=generator))
        fy = float(torch.randint(0, 4, (1,), generator=generator))
        ph = float(torch.rand(1, generator=generator)) * 2 * math.pi
        amp = float(torch.randn(1, generator=generator))
        value = value + amp * torch.cos(2 * math.pi * (fx * x + fy * y) + ph)
    value = (value - value.min()) / (value.max() - value.min())
    return value


def golden_dartel_step() -> None:
    g = torch.Generator().manual_seed(4242)
    out = {}
    for name, (nx, ny) in {"a": (32, 24), "b": (9, 7)}.items():
        src = _smooth_maps(nx, ny, g)
        tgt = torch.roll(src, shifts=(1, 2), dims=(0, 1)) * 0.9 + 0.05
        v = torch.randn((2, ny, nx), dtype=F64, generator=g) * 0.3
        out[f"{name}_src"] = src
        out[f"{name}_tgt"] = tgt
        out[f"{name}_v"] = v
        params = [1.0, 1.0, 0.125, 0.02, 0.001]
        for code in (0, 1):
            for k in range(0, 6):
                cycles, nit = (3, 3) if name == "a" else (1, 1)
                updated, metrics = dg.dartel_step(
                    src, tgt, v, k=k, params=params, lmreg=0.001, cycles=cycles,
                    nit=nit, code=code, kernel="torch", device="cpu",
                )
                out[f"{name}_c{code}_k{k}_v"] = updated
                out[f"{name}_c{code}_k{k}_m"] = metrics
        zero = torch.zeros_like(v)
        updated, metrics = dg.dartel_step(
            src, tgt, zero, k=2, params=params, lmreg=0.001, code=1,
            kernel="torch", device="cpu", assume_initial_nonzero=True,
        )
        out[f"{name}_zero_assume_v"] = updated
        out[f"{name}_zero_assume_m"] = metrics
        updated, metrics = dg.dartel_step(
            src, tgt, zero, k=2, params=params, lmreg=0.001, code=1,
            kernel="torch", device="cpu",
        )
        out[f"{name}_zero_v"] = updated
        out[f"{name}_zero_m"] = metrics
    _save("dartel_step", **out)


def golden_solve() -> None:
    g = torch.Generator().manual_seed(77)
    out = {}
    params = sw.default_dartel_parameters()
    out["params"] = torch.tensor(params, dtype=F64)
    for name, (nx, ny, steps) in {"a": (32, 16, 2), "b": (64, 32, 1)}.items():
        src = torch.stack([_smooth_maps(nx, ny, g) for _ in range(steps)])
        tgt = torch.stack(
            [torch.roll(s, shifts=(1, -2), dims=(0, 1)) for s in src]
        )
        result = sw.solve_dartel_maps(
            src, tgt, loop=6, its=3, code=1, kernel="torch", device="cpu",
        )
        out[f"{name}_src"] = src
        out[f"{name}_tgt"] = tgt
        out[f"{name}_flow"] = result.flow
        out[f"{name}_metrics"] = result.metrics
        result = sw.solve_dartel_maps(
            src, tgt, loop=6, its=3, code=1, kernel="torch", device="cpu",
            assume_initial_nonzero=True,
        )
        out[f"{name}_flow_assume"] = result.flow
        out[f"{name}_metrics_assume"] = result.metrics
    _save("solve", **out)


def golden_warp() -> None:
    g = torch.Generator().manual_seed(5150)
    out = {}
    verts, _faces = _icosphere(3)
    pts = torch.as_tensor(verts, dtype=torch.float32)
    nx, ny = 32, 16
    ident = dg.make_identity_map(dg.GridSpec(nx=nx, ny=ny), device="cpu")
    v = torch.stack([_smooth_maps(nx, ny, g) - 0.5, _smooth_maps(nx, ny, g) - 0.5]) * 3.0
    flow = dg.expdef(v, k=6, device="cpu")
    out["points"] = pts
    out["flow"] = flow
    out["ident"] = ident
    out["warp_inverse"] = sw.apply_flow_to_sphere(pts, flow, inverse=True, device="cpu")
    out["warp_forward"] = sw.apply_flow_to_sphere(pts, flow, inverse=False, device="cpu")
    unit = pts * 1.0
    out["warp_cached"] = sw.apply_flow_to_sphere(
        pts * 3.0, flow, unit_sphere_vertices=unit, inverse=True, device="cpu"
    )
    wraps = torch.linspace(-3.2, 3.2, 129, dtype=F64)
    out["wrap_in"] = wraps
    out["wrap_out"] = sw._wrap_displacement(wraps)
    x_surface = out["warp_inverse"]
    rot = gp._rotation_matrix_from_angles(np.asarray((0.0, math.pi / 2.0, 0.0)))
    z_surface = (out["warp_forward"].to(F64) @ rot.T).to(torch.float32)
    out["avg_x"] = x_surface
    out["avg_z"] = z_surface
    out["avg_out"] = gp._average_xz_surfaces(x_surface, z_surface, device=torch.device("cpu"))
    angles = np.asarray(((0.0, 0.0, 0.0), (0.3, -0.2, 0.1), (-0.6, 0.3, -0.3), (0.0, math.pi / 2, 0.0)))
    out["rot_angles"] = angles
    out["rot_mats"] = np.stack([gp._rotation_matrix_from_angles(a).numpy() for a in angles])
    raw = (verts * 1.3 + np.asarray((0.1, -0.2, 0.05))).astype(np.float32)
    out["norm_in"] = raw
    out["norm_out"] = gp._normalise_sphere(raw)
    out["rotv_out"] = gp._rotate_vertices(out["norm_out"], angles[1])
    _save("warp", **out)


def _small_stencil(verts, faces, nx, ny, generator):
    """A stencil whose indices/weights are random but valid (test fixture)."""

    n = verts.shape[0]
    r = torch.Generator().manual_seed(11)
    surf_idx = faces[torch.randint(0, faces.shape[0], (n,), generator=r).numpy()]
    w = torch.rand((n, 3), dtype=F64, generator=r).numpy()
    w /= w.sum(1, keepdims=True)
    sheet_idx = faces[torch.randint(0, faces.shape[0], (ny * nx,), generator=r).numpy()]
    sw_ = torch.rand((ny * nx, 3), dtype=F64, generator=r).numpy()
    sw_ /= sw_.sum(1, keepdims=True)
    return ss.SurfaceStencil(
        sphere_points=verts.astype(np.float64),
        faces=faces.astype(np.int32),
        surface_indices=surf_idx.astype(np.int32),
        surface_weights=w,
        sheet_indices=sheet_idx.reshape(ny, nx, 3).astype(np.int32),
        sheet_weights=sw_.reshape(ny, nx, 3),
        nx=nx,
        ny=ny,
        source_points=n,
        unit_sphere_points=verts.astype(np.float32),
    )


def golden_surface() -> None:
    g = torch.Generator().manual_seed(31337)
    verts, faces = _icosphere(3)
    nx, ny = 32, 16
    stencil = _small_stencil(verts, faces, nx, ny, g)
    dev = stencil.to("cpu")
    bump = 1.0 + 0.08 * torch.randn(verts.shape[0], dtype=F64, generator=g)
    surface = torch.as_tensor(verts * 30.0, dtype=F64) * bump[:, None]
    surface = surface.to(torch.float32)
    out = {
        "verts": verts, "faces": faces, "surface": surface,
        "surf_idx": stencil.surface_indices, "surf_w": stencil.surface_weights,
        "sheet_idx": stencil.sheet_indices, "sheet_w": stencil.sheet_weights,
    }
    out["resampled"] = dev.resample_vertices(surface)
    out["resampled64"] = dev.resample_vertices(surface.to(F64))
    edge = dev._edge_mean_distance(surface)
    out["edge_mean"] = edge
    it, sigma = dev._heat_parameters(surface, 50.0)
    out["heat50_it"] = np.asarray(it)
    out["heat50_sigma"] = sigma
    out["smooth_geom50"] = dev.smooth_geometry(surface, 50.0)
    out["smooth_geom15"] = dev.smooth_geometry(surface, 15.0)
    out["normals"] = dev.vertex_normals(surface)
    vals = torch.randn(verts.shape[0], dtype=F64, generator=g)
    out["values"] = vals
    out["smooth_vals5"] = dev.smooth_values(vals, surface, 5.0)
    many = dev.smooth_values_many(vals, surface, (5.0, 5.0 / 3.0))
    out["smooth_many0"] = many[0]
    out["smooth_many1"] = many[1]
    out["sheet_of_values"] = dev.map_values_to_sheet(vals)
    fwhms = (5.0 / 9.0, 5.0 / 27.0)
    maps = dev.curvature_type5_to_sheet_many(surface, fwhms)
    out["type5_sheet0"] = maps[0]
    out["type5_sheet1"] = maps[1]
    out["type5_values"] = dev.curvature_type5(surface, 5.0)
    out["type2_values"] = dev.curvature_type2(surface, 5.0)
    out["type2_sheet"] = dev.curvature_to_sheet(surface, 5.0 / 81.0, curvtype=2)
    out["type5_sheet_single"] = dev.curvature_to_sheet(surface, 5.0 / 9.0, curvtype=5)
    _save("surface", **out)


def golden_rotation() -> None:
    g = torch.Generator().manual_seed(2718)
    verts, faces = _icosphere(3)
    index_cpu = sr.RotationGridIndex.from_geometry(verts, faces, grid_size=32, margin=1)
    index = index_cpu.to("cpu")
    src_pts, _ = _icosphere(2)
    src_pts = src_pts.astype(np.float32)
    tv = torch.randn(verts.shape[0], dtype=F64, generator=g)
    # a smooth field: low-order polynomial of the coordinates
    t_pts = torch.as_tensor(verts, dtype=F64)
    tgt_vals = (t_pts[:, 0] * 0.7 + t_pts[:, 1] ** 2 - 0.4 * t_pts[:, 2] * t_pts[:, 0]) + 0.01 * tv
    true_angle = np.asarray((0.12, -0.08, 0.2))
    rot = gp._rotation_matrix_from_angles(true_angle)
    s_pts = torch.as_tensor(src_pts, dtype=F64) @ rot.T
    src_vals = s_pts[:, 0] * 0.7 + s_pts[:, 1] ** 2 - 0.4 * s_pts[:, 2] * s_pts[:, 0]
    inputs = index.prepare_cost_inputs(src_pts, src_vals, tgt_vals)
    seeds = sr.official_seed_grid_angles()
    seed_costs = index.cost_batch_prepared(inputs, torch.as_tensor(seeds), point_chunk=97)
    angle, cost, seed_all = index.coarse_seed_search_prepared(inputs, point_chunk=97)
    ref_angle, ref_cost, iters = index.refine_nelder_mead_prepared(
        inputs, angle, point_chunk=97, max_iter=60, tol=1e-4, simplex_step=0.1
    )
    rotated = index.rotate_points(inputs.source_points, torch.as_tensor(seeds[:5]))
    faces_loc, bary = index._locate_chunk(rotated[:, :200])
    out = {
        "verts": verts, "faces": faces, "src_pts": src_pts,
        "src_vals": src_vals, "tgt_vals": tgt_vals,
        "cand_offsets": index_cpu.candidate_offsets,
        "cand_faces": index_cpu.candidate_faces,
        "seeds": seeds, "seed_costs": seed_costs,
        "seed_angle": angle, "seed_cost": cost,
        "nm_angle": ref_angle, "nm_cost": ref_cost, "nm_iters": np.asarray(iters),
        "rotated5": rotated, "loc_faces": faces_loc, "loc_bary": bary,
        "rotmat": index.rotation_matrix(torch.as_tensor(true_angle)),
    }
    _save("rotation", **out)


def golden_rotation_feature() -> None:
    """The 'cuda' rotation-feature backend on a small stencil (rotation_feature.py:320-399)."""

    from cat_surface_gpu import rotation_feature as rf

    g = torch.Generator().manual_seed(1618)
    verts, faces = _icosphere(3)
    nx, ny = 32, 16
    stencil = _small_stencil(verts, faces, nx, ny, g)
    dev = stencil.to("cpu")
    bump = 1.0 + 0.08 * torch.randn(verts.shape[0], dtype=F64, generator=g)
    surface = (torch.as_tensor(verts * 30.0, dtype=F64) * bump[:, None]).to(torch.float32)
    smoothed = dev.smooth_geometry(dev.resample_vertices(surface), 15.0)
    normals = rf._stable_normals(smoothed, dev.faces)
    areas = rf._mixed_voronoi_areas(smoothed, dev.faces)
    depth, iters, resid = rf._solve_depth_potential_colored_sor(
        smoothed, dev.faces, areas, normals, dev.color_groups(), alpha=1e-3, max_iter=1000)
    result = rf.compute_rotation_feature(dev, surface, heat_fwhm=15.0)
    depth_cg, cg_iters, cg_resid = rf._solve_depth_potential_cg(
        smoothed, dev.faces, areas, normals, alpha=1e-3, max_iter=1000, tolerance=1e-9)
    result_cg = rf.compute_rotation_feature(dev, surface, heat_fwhm=15.0, solver="pcg")
    groups = dev.color_groups()
    colours = np.zeros(verts.shape[0], np.int32)
    for c, grp in enumerate(groups):
        colours[grp.numpy()] = c
    _save("rotation_feature", verts=verts, faces=faces, surface=surface,
          surf_idx=stencil.surface_indices, surf_w=stencil.surface_weights,
          smoothed=smoothed, normals=normals, areas=areas, depth=depth,
          residual=np.asarray(resid), colours=colours, feature=result.values,
          depth_cg=depth_cg, cg_iterations=np.asarray(cg_iters),
          cg_residual=np.asarray(cg_resid), feature_cg=result_cg.values)


def main() -> None:
    steps = [
        golden_grid_ops, golden_objfun_squaring, golden_solver,
        golden_dartel_step, golden_solve, golden_warp, golden_surface,
        golden_rotation, golden_rotation_feature,
    ]
    only = set(sys.argv[1:])
    for step in steps:
        if only and step.__name__ not in only:
            continue
        print(f"-- {step.__name__}", flush=True)
        step()


if __name__ == "__main__":
    main()
