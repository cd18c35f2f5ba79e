"""gr_intersect_closest_noalloc against gr_intersect_closest, through the C ABI.

The no-allocation variant keeps the ray-sort scratch and the traversal's work
counters in one static device array per GPU; its hits and per-ray work counts
must equal the allocating entry point's bit for bit, on two calls in a row (the
second with fewer rays, so stale scratch from the first call is present), with
the ray sort on and off.
"""
import ctypes

import numpy as np
import pytest

from gammaray_amd import _lib, scenes
from gammaray_amd.api import Stage

pytestmark = pytest.mark.gpu


def _run(stage, rays, noalloc):
    import torch
    n = rays.shape[0]
    hits = torch.full((n, 4), -7, dtype=torch.int32, device=rays.device)
    cnt = torch.full((n, 3), -7, dtype=torch.int32, device=rays.device)
    if noalloc:
        stage.intersect_noalloc_ptr(rays.data_ptr(), n, hits.data_ptr(), cnt.data_ptr())
    else:
        stage.intersect_ptr(rays.data_ptr(), n, hits.data_ptr(), cnt.data_ptr())
    torch.cuda.synchronize()
    return hits.cpu().numpy(), cnt.cpu().numpy()


@pytest.mark.parametrize("sort_bits,xcd", [(15, 0), (0, 1)])
def test_noalloc_matches_allocating_entry_point(gpu_ctx, sort_bits, xcd):
    import torch
    spec = scenes.icosphere_scene(level=6)
    stage = Stage(spec, ctx=gpu_ctx)
    stage.set_tuning(sort_bits=sort_bits, xcd_ranges=xcd)
    big = _lib.GR_NOALLOC_MAX_RAYS
    rays = torch.from_numpy(scenes.ray_batch(spec, big, seed=11)).cuda()
    for n in (big, 100_003):
        print("noalloc case: sort_bits=%d xcd=%d n=%d" % (sort_bits, xcd, n), flush=True)
        r = rays[:n].contiguous()
        want_h, want_c = _run(stage, r, noalloc=False)
        got_h, got_c = _run(stage, r, noalloc=True)
        assert np.array_equal(got_h, want_h)
        assert np.array_equal(got_c, want_c)
        assert (want_h[:, 0] != -7).all()
    stage.close()


def test_noalloc_rejects_more_than_the_maximum(gpu_ctx):
    spec = scenes.box_scene()
    stage = Stage(spec, ctx=gpu_ctx)
    code = _lib.lib().gr_intersect_closest_noalloc(stage.handle, ctypes.c_void_p(16), _lib.GR_NOALLOC_MAX_RAYS + 1,
                                                   ctypes.c_void_p(16), None)
    assert code == -1
    assert b"GR_NOALLOC_MAX_RAYS" in _lib.lib().gr_ctx_last_error(gpu_ctx.handle)
    stage.close()
