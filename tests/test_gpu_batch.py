"""Batched rendering through one shared pool of in-flight paths: every job's
film must equal the film gr_trace renders for that job alone -- bit for bit
-- whatever the pool size (paths are regenerated into freed slots many times
over), with the bounce step replayed as a hipGraph or launched eagerly, with
sorted multi-scene traversal, for PT and PT+NEE."""
import numpy as np
import pytest

from gammaray_amd import scenes
from gammaray_amd.api import Film, Stage, render_jobs, trace

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def stages(gpu_ctx):
    box = Stage(scenes.box_scene(), ctx=gpu_ctx)
    show = Stage(scenes.showcase_scene(n_spheres=60, grid=(40, 40)), ctx=gpu_ctx)
    room = Stage(scenes.indirect_room_scene(), ctx=gpu_ctx)
    return box, show, room


def _jobs(ctx, stages):
    box, show, room = stages
    other = scenes.CameraSpec(focal_length=1.2, horizontal_aperture=1.6, vertical_aperture=0.9,
                              rotate=(0.9659258, 0.0, 0.258819, 0.0), translate=(0.5, 1.4, 3.0))
    return [
        dict(stage=box, film=Film(ctx, 40, 32), spp=5, max_depth=6, seed=3),
        dict(stage=show, film=Film(ctx, 48, 27), spp=3, max_depth=8, seed=4, sample_offset=2),
        dict(stage=room, film=Film(ctx, 32, 32), spp=7, max_depth=0, seed=5),
        dict(stage=box, film=Film(ctx, 17, 23), spp=2, max_depth=3, seed=6),
        dict(stage=show, film=Film(ctx, 64, 36), spp=1, max_depth=8, seed=7, camera=other),
    ]


def _alone(ctx, job, integrator):
    f = Film(ctx, job["film"].width, job["film"].height)
    trace(job["stage"], f, integrator, spp=job["spp"], max_depth=job["max_depth"], seed=job["seed"],
          sample_offset=job.get("sample_offset", 0), camera=job.get("camera"))
    return f.readback()


@pytest.mark.parametrize("graph", [True, False])
@pytest.mark.parametrize("integrator", ["pt_nee", "pt"])
def test_batch_films_equal_single_job_renders(graph, integrator, gpu_ctx, stages):
    print("case batch", integrator, "graph" if graph else "eager", flush=True)
    jobs = _jobs(gpu_ctx, stages)
    ms = render_jobs(gpu_ctx, jobs, integrator, path_pool=3000, steps_per_poll=8, use_graph=graph)
    assert all(np.isfinite(m) and m > 0 for m in ms)
    for k, job in enumerate(jobs):
        got = job["film"].readback()
        want = _alone(gpu_ctx, job, integrator)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), k
        assert np.abs(got[..., :3]).sum() > 0


def test_batch_with_sorted_multi_scene_traversal(gpu_ctx, stages):
    print("case batch sorted", flush=True)
    jobs = _jobs(gpu_ctx, stages)
    render_jobs(gpu_ctx, jobs, "pt_nee", path_pool=20000, steps_per_poll=4,
                tuning=dict(sort_bits=18, lds_stack=8, refill=32))
    for k, job in enumerate(jobs):
        assert np.array_equal(job["film"].readback().view(np.uint32),
                              _alone(gpu_ctx, job, "pt_nee").view(np.uint32)), k
