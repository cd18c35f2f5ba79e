"""The `gammaray` driver (the reference binary's progressive loop,
src/bin/gammaray.rs:81-112): per-pass lines, the EXR rewritten after every
pass, and -- with its readback overlapped with rendering -- a file byte-equal
to the same passes rendered through the API and written by gr_film_write_exr."""
import os
import subprocess

import numpy as np
import pytest

from gammaray_amd import scenes
from gammaray_amd.api import Film, Stage, trace

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CLI = os.path.join(HERE, "..", "gammaray_amd", "gammaray")
OBJ = os.path.join(HERE, "golden", "obj", "sphere.obj")


@pytest.mark.parametrize("integrator,iters", [("bdpt", 4), ("pt_nee", 3), ("bdpt_mis", 3), ("bdpt_lt", 3)])
def test_driver_output_equals_api_render(integrator, iters, gpu_ctx, tmp_path):
    print("case cli", integrator, flush=True)
    out = str(tmp_path / "out.exr")
    r = subprocess.run([CLI, "--iterations", str(iters), "--height", "64", "--integrator", integrator,
                        "--obj", OBJ, "--output", out, "--max-depth", "6"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[0].startswith("Aspect ratio: 1.375, Width: 88, Height: 64")
    assert [l.split()[1] for l in lines[1:]] == [str(k) for k in range(iters)]
    spec = scenes.demo_scene(OBJ)
    stage = Stage(spec, ctx=gpu_ctx)
    film = Film(gpu_ctx, 88, 64)
    # one iteration per batch, as the driver's passes (light tracing adds each
    # batch's splats to the film as one estimate)
    trace(stage, film, integrator, spp=iters, max_depth=6, seed=1, batch=1)
    ref = str(tmp_path / "ref.exr")
    film.write_exr(ref)
    a, b = open(out, "rb").read(), open(ref, "rb").read()
    assert a == b
    px = film.readback()
    assert np.isfinite(px).all() and px[..., :3].sum() > 0


def test_driver_rejects_a_missing_required_mesh(tmp_path):
    r = subprocess.run([CLI, "--iterations", "1", "--height", "8", "--obj", str(tmp_path / "none.obj"),
                        "--output", str(tmp_path / "o.exr")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "Couldn't open OBJ file" in r.stderr


def test_overlapped_readback_is_the_film_at_its_point_in_the_queue(gpu_ctx):
    """readback_begin after pass k, then pass k+1 queued: readback_end returns
    the film after pass k, not after k+1."""
    print("case readback overlap", flush=True)
    spec = scenes.box_scene()
    stage = Stage(spec, ctx=gpu_ctx)
    f = Film(gpu_ctx, 48, 40)
    trace(stage, f, "pt", spp=2, max_depth=4, seed=3)
    want = f.readback()
    f.readback_begin()
    trace(stage, f, "pt", spp=2, max_depth=4, seed=3, sample_offset=2)
    got = f.readback_end()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert not np.array_equal(f.readback(), want)
    with pytest.raises(Exception):
        f.readback_end()  # nothing in flight

