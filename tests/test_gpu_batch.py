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


@pytest.mark.parametrize("ring,order", [(1, "as_given"), (5000, "fewest_first"), (0, "fewest_first")])
def test_batch_sample_ring_and_job_order(gpu_ctx, stages, ring, order):
    """Finished samples wait in a ring before each job's iterations are
    gathered in order.  A ring smaller than the pool (raised to the largest
    film, 64 x 36 pixels) wraps many times and holds generation back; the
    films must still equal gr_trace's bit for bit, whatever order the jobs
    take samples in."""
    print("case batch ring", ring, order, flush=True)
    jobs = _jobs(gpu_ctx, stages)
    ms = render_jobs(gpu_ctx, jobs, "pt_nee", path_pool=3000, steps_per_poll=2, sample_ring=ring,
                     job_order=order)
    assert all(np.isfinite(m) and m > 0 for m in ms)
    for k, job in enumerate(jobs):
        assert np.array_equal(job["film"].readback().view(np.uint32),
                              _alone(gpu_ctx, job, "pt_nee").view(np.uint32)), k
    if order == "fewest_first":
        # the job with the fewest samples finishes first
        samples = [j["film"].width * j["film"].height * j["spp"] for j in jobs]
        assert int(np.argmin(ms)) == int(np.argmin(samples))


def test_batch_rejects_unknown_job_order(gpu_ctx, stages):
    from gammaray_amd import _lib
    from gammaray_amd.api import JOB_ORDERS
    print("case batch order validation", flush=True)
    JOB_ORDERS["bogus"] = 7
    try:
        with pytest.raises(_lib.GrError):
            render_jobs(gpu_ctx, _jobs(gpu_ctx, stages)[:1], "pt_nee", job_order="bogus")
    finally:
        del JOB_ORDERS["bogus"]


def test_full_size_batched_workload(gpu_ctx):
    """The 64-job configuration at full size: 8 displaced-terrain scenes of
    10k..2M triangles, films 128^2..1024^2, spp 4..256, PT+NEE depth 8, 1.43G
    samples through the default pool and sample ring, bounce step as a graph.
    Four jobs (the smallest and largest film, the smallest and largest scene)
    equal gr_trace's films bit for bit; the smallest job also matches the CPU
    oracle (film weights bit-exact, per-sample radiance paired)."""
    import torch
    from oracle import oracle as O
    print("case full batch: building 8 scenes", flush=True)
    specs, jobs = scenes.batch_workload(seed=4)
    stages = [Stage(sp, ctx=gpu_ctx) for sp in specs]
    films = [Film(gpu_ctx, j["side"], j["side"]) for j in jobs]
    jl = [dict(stage=stages[j["scene"]], film=films[k], spp=j["spp"], max_depth=8, seed=j["seed"])
          for k, j in enumerate(jobs)]
    print("case full batch: rendering %d samples" % sum(j["side"] ** 2 * j["spp"] for j in jobs), flush=True)
    ms = render_jobs(gpu_ctx, jl, "pt_nee", use_graph=True)
    assert all(np.isfinite(m) and m > 0 for m in ms)
    side = [j["side"] for j in jobs]
    tris = [specs[j["scene"]].num_components for j in jobs]
    picks = sorted({int(np.argmin(side)), int(np.argmax(side)), int(np.argmin(tris)), int(np.argmax(tris))})
    for k in picks:
        job = jl[k]
        print("case full batch: job %d (%d^2, %d spp, %d triangles) vs gr_trace"
              % (k, side[k], job["spp"], tris[k]), flush=True)
        want = _alone(gpu_ctx, job, "pt_nee")
        assert np.array_equal(job["film"].readback().view(np.uint32), want.view(np.uint32)), k
    # the cheapest job against the oracle
    k = int(np.argmin([side[i] ** 2 * jobs[i]["spp"] * tris[i] for i in range(len(jobs))]))
    job, spec = jl[k], specs[jobs[k]["scene"]]
    print("case full batch: job %d vs oracle" % k, flush=True)
    W = H = side[k]
    f = Film(gpu_ctx, W, H)
    buf = torch.zeros((job["spp"] * W * H, 4), dtype=torch.float32, device="cuda")
    trace(job["stage"], f, "pt_nee", spp=job["spp"], max_depth=8, seed=job["seed"], sample_radiance_ptr=buf.data_ptr())
    gpu_ctx.synchronize()
    g = buf.cpu().numpy()[:, :3].reshape(job["spp"], W * H, 3)
    o = O.OracleScene(spec)
    ref, r = O.render(o, spec, W, H, O.PT_NEE, spp=job["spp"], max_depth=8, seed=job["seed"], colors=True,
                      threads=16)
    assert np.array_equal(job["film"].readback()[:, 3].view(np.uint32),
                          np.asarray(ref, np.float32).reshape(-1, 4)[:, 3].view(np.uint32))
    close = np.all(np.isclose(g, r, rtol=1e-3, atol=1e-5), axis=2)
    assert close.mean() > 0.99, close.mean()
    np.testing.assert_allclose(g.reshape(-1, 3).mean(0), r.reshape(-1, 3).mean(0), rtol=2e-2)
    for fl in films:
        fl.close()
    for st in stages:
        st.close()
