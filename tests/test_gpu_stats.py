"""GPU: the context's device-side error counters (gr_ctx_take_stats).

The reference asserts, in debug builds, that every material sample has a
finite radiance and a finite positive pdf (src/material/material.rs:224-226),
and lets non-finite values through in release builds.  The library keeps the
release behaviour (the values reach the film) and counts each event:
  * nonfinite_bsdf_samples: material samples failing those assertions;
  * nonfinite_radiance: camera samples whose radiance reaches the film
    non-finite -- exactly the non-finite rows of the per-sample export;
  * bad_splats: light-tracing contributions dropped (non-finite) or clamped.
A scene with an infinite base colour raises them; an ordinary scene leaves
them at zero.
"""
import numpy as np
import pytest

from gammaray_amd import scenes
from gammaray_amd.api import Film, Stage, trace

pytestmark = pytest.mark.gpu


def _poisoned_box():
    spec = scenes.box_scene(seed=1)
    # the first sphere gets an infinite base colour
    for p in spec.prims:
        if p.kind == "sphere":
            p.material = p.material.copy()
            p.material[0] = 0.0
            p.material[1:4] = np.inf
            break
    return spec


def _render(ctx, spec, integrator, spp, **kw):
    import torch
    W, H = 48, 48
    stage = Stage(spec, ctx=ctx)
    film = Film(ctx, W, H)
    buf = torch.zeros((spp * W * H, 4), dtype=torch.float32, device="cuda")
    ctx.take_stats()
    trace(stage, film, integrator, spp=spp, sample_radiance_ptr=buf.data_ptr(), seed=5, **kw)
    st = ctx.take_stats()
    rad = buf.cpu().numpy()[:, :3]
    film.close()
    stage.close()
    return st, rad


@pytest.mark.parametrize("integrator", ["pt", "pt_nee", "bdpt", "bdpt_lt"])
def test_clean_scene_counts_nothing(gpu_ctx, integrator):
    print("case stats clean", integrator, flush=True)
    st, rad = _render(gpu_ctx, scenes.box_scene(seed=1), integrator, 4, max_depth=6, light_depth=3)
    assert np.isfinite(rad).all()
    assert st == {"nonfinite_bsdf_samples": 0, "nonfinite_radiance": 0, "bad_splats": 0}


@pytest.mark.parametrize("integrator", ["pt", "pt_nee", "bdpt", "bdpt_lt"])
def test_infinite_base_colour_is_counted(gpu_ctx, integrator):
    print("case stats poisoned", integrator, flush=True)
    st, rad = _render(gpu_ctx, _poisoned_box(), integrator, 4, max_depth=6, light_depth=3)
    bad_rows = int((~np.isfinite(rad)).any(axis=1).sum())
    assert bad_rows > 0
    assert st["nonfinite_bsdf_samples"] > 0
    # every non-finite camera sample is counted once, when it reaches the film
    assert st["nonfinite_radiance"] == bad_rows
    if integrator == "bdpt_lt":
        assert st["bad_splats"] > 0
    else:
        assert st["bad_splats"] == 0
    # the counters were cleared by the read
    assert gpu_ctx.take_stats() == {"nonfinite_bsdf_samples": 0, "nonfinite_radiance": 0, "bad_splats": 0}


def test_objects_may_outlive_their_context():
    """Scenes, films and guides keep their context alive; destroying the
    context's handle first, then the objects, is safe (as happens when a
    garbage collector runs after a context was closed)."""
    from gammaray_amd.api import Context, Guide
    print("case context lifetime", flush=True)
    ctx = Context(0)
    stage = Stage(scenes.box_scene(), ctx=ctx)
    film = Film(ctx, 16, 16)
    guide = Guide(ctx, batch=1024, train_paths=1024)
    trace(stage, film, "pt", spp=2, max_depth=4)
    ctx.close()
    img = film.readback()
    assert np.isfinite(img).all() and img[:, 3].sum() > 0
    guide.close()
    stage.close()
    film.close()
