"""End-to-end CAT_SurfWarp pipeline on the GPU, on a small synthetic pair.

The reference CPU CAT_SurfWarp binary is not runnable here (it ships as a
prebuilt executable), so end-to-end parity is checked through invariants:
exact triangles, unit-length output, run-to-run bit identity, and a rotation
search that lowers the feature SSD well below the identity's.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pair_paths(tmp_path_factory):
    from cat_surface_gpu import synthetic

    pair = synthetic.make_pair("icosphere", subdivisions=4, lmax=6, noise=0.0,
                               max_rotation_deg=15.0, seed=99)
    folder = tmp_path_factory.mktemp("pair")
    return pair, synthetic.write_pair(pair, folder), folder


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_pipeline_end_to_end(pair_paths, cuda, dtype):
    from cat_surface_gpu import read_gifti_mesh, run_cat_surface_gpu_pipeline

    pair, paths, folder = pair_paths
    out = folder / f"out_{dtype}.gii"
    res = run_cat_surface_gpu_pipeline(paths["src_surface"], paths["src_sphere"],
                                       paths["tgt_surface"], paths["tgt_sphere"], out,
                                       device=cuda, dartel_dtype=dtype, steps=2, runs=2,
                                       avg=True)
    mesh = read_gifti_mesh(out)
    np.testing.assert_array_equal(mesh.faces, pair.src_faces)
    np.testing.assert_allclose(np.linalg.norm(mesh.vertices, axis=1), 1.0, atol=1e-5)
    assert np.all(np.isfinite(mesh.vertices))
    again = run_cat_surface_gpu_pipeline(paths["src_surface"], paths["src_sphere"],
                                         paths["tgt_surface"], paths["tgt_sphere"], out,
                                         device=cuda, dartel_dtype=dtype, steps=2, runs=2,
                                         avg=True)
    np.testing.assert_array_equal(again.vertices, res.vertices)
    assert res.rotation_cost >= 0.0
    for key in ("rotation_search_seconds", "dartel_total_seconds", "total_seconds"):
        assert key in res.timings


def test_batched_jobs_equal_single_runs(pair_paths, cuda, tmp_path):
    """Two registrations scheduled together give exactly the single-job results."""

    from cat_surface_gpu import (GiftiMesh, PipelineJob, run_cat_surface_gpu_pipeline,
                                 run_cat_surface_gpu_pipelines, synthetic)

    pair, paths, folder = pair_paths
    mirrored = synthetic.make_pair("icosphere", subdivisions=4, lmax=6, noise=0.0,
                                   max_rotation_deg=15.0, seed=99, mirror=True)
    jobs = []
    for p in (pair, mirrored):
        jobs.append(PipelineJob(meshes=(GiftiMesh(p.src_surface, p.src_faces),
                                        GiftiMesh(p.src_sphere, p.src_faces),
                                        GiftiMesh(p.tgt_surface, p.tgt_faces),
                                        GiftiMesh(p.tgt_sphere, p.tgt_faces))))
    kw = dict(device=cuda, dartel_dtype="float32", steps=2, runs=2, avg=True,
              write_output=False)
    together = run_cat_surface_gpu_pipelines(jobs, **kw)
    for job, res in zip(jobs, together):
        alone = run_cat_surface_gpu_pipeline(None, None, None, None, tmp_path / "x.gii",
                                             meshes=job.meshes, **kw)
        np.testing.assert_array_equal(res.vertices, alone.vertices)
        np.testing.assert_array_equal(res.rotation_angle, alone.rotation_angle)
        np.testing.assert_array_equal(res.faces, alone.faces)


@pytest.mark.parametrize("graph_schedule", [False, True])
def test_device_resident_meshes_equal_host_arrays(pair_paths, cuda, graph_schedule):
    """Meshes handed over as device tensors (nothing copied from the host) give
    the host arrays' results bit for bit; the result's faces are the tensor given."""

    from cat_surface_gpu import GiftiMesh, PipelineJob, run_cat_surface_gpu_pipelines
    from cat_surface_gpu import gpu_pipeline as gp

    pair, paths, folder = pair_paths
    host = (GiftiMesh(pair.src_surface, pair.src_faces), GiftiMesh(pair.src_sphere, pair.src_faces),
            GiftiMesh(pair.tgt_surface, pair.tgt_faces), GiftiMesh(pair.tgt_sphere, pair.tgt_faces))
    resident = tuple(GiftiMesh(torch.as_tensor(m.vertices, device=cuda),
                               torch.as_tensor(m.faces, device=cuda)) for m in host)
    kw = dict(device=cuda, dartel_dtype="float32", steps=2, runs=2, avg=True,
              write_output=False, rotation_method="adjoint", graph_schedule=graph_schedule)
    gp.clear_graph_schedules()
    ref = run_cat_surface_gpu_pipelines([PipelineJob(meshes=host)], **kw)[0]
    for _ in range(2):  # with graph_schedule: capture, then replay
        got = run_cat_surface_gpu_pipelines([PipelineJob(meshes=resident)], **kw)[0]
        np.testing.assert_array_equal(got.vertices, ref.vertices)
        np.testing.assert_array_equal(np.asarray(got.rotation_angle),
                                      np.asarray(ref.rotation_angle))
        assert got.faces is resident[1].faces
    if graph_schedule:
        assert got.timings.get("graph_replay") == 1.0


def test_adjoint_rotation_pipeline_matches_nelder_mead(pair_paths, cuda):
    """Gradient-refined rotation: no higher SSD than Nelder-Mead's choice (1e-6
    relative) and a registration within the 1e-3 vertex tolerance of it."""

    from cat_surface_gpu import GiftiMesh, PipelineJob, run_cat_surface_gpu_pipelines

    pair, _, _ = pair_paths
    job = PipelineJob(meshes=(GiftiMesh(pair.src_surface, pair.src_faces),
                              GiftiMesh(pair.src_sphere, pair.src_faces),
                              GiftiMesh(pair.tgt_surface, pair.tgt_faces),
                              GiftiMesh(pair.tgt_sphere, pair.tgt_faces)))
    kw = dict(device=cuda, dartel_dtype="float32", steps=2, runs=2, avg=True,
              write_output=False)
    nm = run_cat_surface_gpu_pipelines([job], **kw)[0]
    for candidates in ("seeds", "hopf"):
        adj = run_cat_surface_gpu_pipelines([job], rotation_method="adjoint",
                                            rotation_candidates=candidates, **kw)[0]
        assert adj.rotation_cost <= nm.rotation_cost * (1.0 + 1e-6)
        np.testing.assert_array_equal(adj.faces, nm.faces)
        assert np.abs(adj.vertices - nm.vertices).max() <= 1e-3


def test_uneven_jobs_equal_single_runs(cuda, tmp_path):
    """Jobs of different mesh and map sizes in one schedule (one batched solve per
    map size, on streams of their own, with tuned tiles) give the single-job results."""

    from cat_surface_gpu import (GiftiMesh, PipelineJob, run_cat_surface_gpu_pipeline,
                                 run_cat_surface_gpu_pipelines, synthetic)

    jobs = []
    for k, (sub, size) in enumerate(((3, (128, 64)), (4, (256, 128)), (3, (256, 128)))):
        p = synthetic.make_pair("icosphere", subdivisions=sub, lmax=6, noise=0.0,
                                max_rotation_deg=15.0, seed=300 + k)
        jobs.append(PipelineJob(meshes=(GiftiMesh(p.src_surface, p.src_faces),
                                        GiftiMesh(p.src_sphere, p.src_faces),
                                        GiftiMesh(p.tgt_surface, p.tgt_faces),
                                        GiftiMesh(p.tgt_sphere, p.tgt_faces)), map_size=size))
    kw = dict(device=cuda, dartel_dtype="float32", steps=2, runs=2, avg=True,
              write_output=False)
    together = run_cat_surface_gpu_pipelines(jobs, autotune=True, **kw)
    for job, res in zip(jobs, together):
        alone = run_cat_surface_gpu_pipelines([job], **kw)[0]
        np.testing.assert_array_equal(res.vertices, alone.vertices)
        np.testing.assert_array_equal(res.faces, alone.faces)
        assert np.all(np.isfinite(res.vertices))
