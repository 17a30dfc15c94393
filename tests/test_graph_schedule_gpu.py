"""The whole batched schedule captured in one HIP graph (run_graph_schedule).

Replays are compared bit for bit with the eager schedule of the same jobs:
the graph holds the same kernels with the data-dependent loop counts read on
the device (heat iterations) and every host decision (rotation choice, sphere
rotation, check of stencil capacities) moved to the device.  A replay whose
counts exceed the recorded bounds falls back to the eager schedule.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _jobs(seeds, scale=1.0, subdivisions=4, map_size=(128, 64)):
    from cat_surface_gpu import GiftiMesh, PipelineJob, synthetic

    jobs = []
    for seed in seeds:
        p = synthetic.make_pair("icosphere", subdivisions=subdivisions, lmax=8, noise=0.0,
                                max_rotation_deg=15.0, warp_deg=3.0, seed=seed)
        meshes = (GiftiMesh(p.src_surface * np.float32(scale), p.src_faces),
                  GiftiMesh(p.src_sphere, p.src_faces),
                  GiftiMesh(p.tgt_surface * np.float32(scale), p.tgt_faces),
                  GiftiMesh(p.tgt_sphere, p.tgt_faces))
        jobs.append(PipelineJob(meshes=meshes, map_size=map_size))
    return jobs


KW = dict(dartel_dtype="float32", steps=2, runs=2, avg=True, write_output=False,
          cuda_graph=True, rotation_method="adjoint")


def _same(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x.vertices, y.vertices)
        np.testing.assert_array_equal(x.faces, y.faces)
        np.testing.assert_array_equal(np.asarray(x.rotation_angle), np.asarray(y.rotation_angle))
        assert float(x.rotation_cost) == float(y.rotation_cost)


def test_graph_schedule_replays_equal_the_eager_schedule(cuda):
    from cat_surface_gpu import gpu_pipeline as gp

    gp.clear_graph_schedules()
    jobs = _jobs((11, 12))
    eager = gp.run_cat_surface_gpu_pipelines(jobs, device=cuda, **KW)
    planned = gp.run_cat_surface_gpu_pipelines(jobs, device=cuda, graph_schedule=True, **KW)
    _same(planned, eager)
    entry = next(iter(gp._GRAPH_SCHEDULES.values()))
    print(f"\ncaptured in {entry.capture_seconds:.2f} s; "
          f"{len(entry.capture.bounds)} bounded loops, {len(entry.capture.flags)} checks",
          flush=True)
    for _ in range(3):
        replayed = gp.run_cat_surface_gpu_pipelines(jobs, device=cuda, graph_schedule=True, **KW)
        assert all(r.timings.get("graph_replay") == 1.0 for r in replayed)
        _same(replayed, eager)
        # eager work of the program between the replays
        gp.run_cat_surface_gpu_pipelines(jobs[:1], device=cuda, **KW)
    # other subjects of the same shapes: the graph serves them too
    other = _jobs((21, 22))
    replayed = gp.run_cat_surface_gpu_pipelines(other, device=cuda, graph_schedule=True, **KW)
    assert all(r.timings.get("graph_replay") == 1.0 for r in replayed)
    _same(replayed, gp.run_cat_surface_gpu_pipelines(other, device=cuda, **KW))
    assert entry.replays == 4 and entry.fallbacks == 0


def test_graph_schedule_with_uneven_map_sizes(cuda):
    """Jobs of two mesh sizes and two map sizes: the solves of each map size run as
    parallel branches of the one graph; replays equal the eager schedule."""

    from cat_surface_gpu import gpu_pipeline as gp

    gp.clear_graph_schedules()
    jobs = _jobs((51,), subdivisions=4, map_size=(128, 64)) + \
        _jobs((52,), subdivisions=5, map_size=(256, 128))
    eager = gp.run_cat_surface_gpu_pipelines(jobs, device=cuda, **KW)
    gp.run_cat_surface_gpu_pipelines(jobs, device=cuda, graph_schedule=True, **KW)
    for _ in range(2):
        replayed = gp.run_cat_surface_gpu_pipelines(jobs, device=cuda, graph_schedule=True, **KW)
        assert all(r.timings.get("graph_replay") == 1.0 for r in replayed)
        _same(replayed, eager)


def test_graph_schedule_falls_back_when_counts_exceed_the_bounds(cuda):
    """Surfaces at half the size have a quarter of the edge length squared, so
    their heat-kernel iteration counts are four times the recorded ones: the
    replay's device check fails and the call runs eagerly (same result as the
    eager schedule)."""

    from cat_surface_gpu import gpu_pipeline as gp

    gp.clear_graph_schedules()
    jobs = _jobs((31,))
    gp.run_cat_surface_gpu_pipelines(jobs, device=cuda, graph_schedule=True, **KW)
    small = _jobs((31,), scale=0.5)
    got = gp.run_cat_surface_gpu_pipelines(small, device=cuda, graph_schedule=True, **KW)
    assert got[0].timings.get("graph_fallback") == 1.0
    entry = next(iter(gp._GRAPH_SCHEDULES.values()))
    assert entry.fallbacks == 1
    assert any("iterations above" in label for label in entry.capture.verdict())
    _same(got, gp.run_cat_surface_gpu_pipelines(small, device=cuda, **KW))


def test_graph_schedule_rejects_host_decided_stages(cuda):
    from cat_surface_gpu import gpu_pipeline as gp

    jobs = _jobs((41,))
    with pytest.raises(ValueError, match="adjoint"):
        gp.run_cat_surface_gpu_pipelines(jobs, device=cuda, graph_schedule=True,
                                         **dict(KW, rotation_method="nelder-mead"))
    with pytest.raises(ValueError, match="type-5"):
        gp.run_cat_surface_gpu_pipelines(jobs, device=cuda, graph_schedule=True,
                                         **dict(KW, steps=3))


def test_stream_of_gifti_batches_equals_the_graph_schedule(cuda, tmp_path):
    """run_cat_surface_gpu_stream (batch k + 1 read and batch k written while the
    GPU registers) gives each batch the results of run_graph_schedule on the
    same meshes, and writes them."""

    from cat_surface_gpu import PipelineJob, gifti, gpu_pipeline as gp, run_cat_surface_gpu_stream

    gp.clear_graph_schedules()
    batches = []
    for b, seeds in enumerate(((11, 12), (21, 22), (31, 32))):
        batch = []
        for seed, job in zip(seeds, _jobs(seeds)):
            paths = [tmp_path / f"b{b}_{seed}_{n}.gii"
                     for n in ("src_surface", "src_sphere", "tgt_surface", "tgt_sphere")]
            for path, mesh in zip(paths, job.meshes):
                gifti.write_gifti_surface(path, mesh.vertices, mesh.faces, compress_level=6)
            batch.append(PipelineJob(*paths, output=tmp_path / f"b{b}_{len(batch)}.sphere.reg.gii",
                                     map_size=job.map_size))
        batches.append(batch)
    streamed = run_cat_surface_gpu_stream(batches, device=cuda, **dict(KW, write_output=True))
    assert len(streamed) == len(batches)
    for batch, results in zip(batches, streamed):
        direct = gp.run_cat_surface_gpu_pipelines(gp._read_job_meshes(batch), device=cuda,
                                                  graph_schedule=True, **KW)
        _same(results, direct)
        for job, r in zip(batch, results):
            written = gifti.read_gifti(job.output)
            np.testing.assert_array_equal(written[0].data, r.vertices)
            np.testing.assert_array_equal(written[1].data, r.faces)
    entry = next(iter(gp._GRAPH_SCHEDULES.values()))
    assert entry.fallbacks == 0
