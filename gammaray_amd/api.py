"""Python host API over the C ABI, named after the reference's types.

``Stage`` is the counterpart of gammaray's ``Stage`` (src/render/stage.rs): it
takes the scene's prims, builds the BVH with the reference's SAH builder and
owns the GPU copy.  ``intersect`` / ``visibility`` are the batched forms of
``Bvh::intersect`` / ``Bvh::visibility`` (src/geom/bvh.rs:406-477) over rays
held in device memory (torch tensors on the GPU are used as plain buffers).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import GR_HOST_ONLY, GR_NONE, GR_OCCLUSION_ANYHIT, GR_OCCLUSION_EXACT, check, lib

_f32p = ctypes.POINTER(ctypes.c_float)
_u32p = ctypes.POINTER(ctypes.c_uint32)


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


class Context:
    """A device context (one HIP stream).  device=GR_HOST_ONLY builds scenes
    without touching a GPU."""

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        code = lib().gr_ctx_create(device, ctypes.byref(self._h))
        if code != 0:
            raise _lib.GrError("gr_ctx_create(%d) failed: %s" % (device, _lib.ERRORS.get(code, code)))
        self.device = device

    @property
    def handle(self):
        return self._h

    def set_stream(self, stream_ptr):
        check(lib().gr_ctx_set_stream(self._h, ctypes.c_void_p(stream_ptr)), self._h, "set_stream")

    def stream(self):
        return lib().gr_ctx_stream(self._h)

    def synchronize(self):
        check(lib().gr_ctx_synchronize(self._h), self._h, "synchronize")

    def take_stats(self):
        """The device-side error counters of this context's render calls since
        the previous call (gr_ctx_take_stats), as a dict; clears them."""
        st = _lib.Stats()
        check(lib().gr_ctx_take_stats(self._h, ctypes.byref(st)), self._h, "take_stats")
        return {k: getattr(st, k) for k, _ in _lib.Stats._fields_ if k != "reserved"}

    KERNEL_KINDS = ("trace_closest", "trace_shadow", "ray_sort")

    def set_kernel_timing(self, enable=True):
        check(lib().gr_ctx_set_kernel_timing(self._h, 1 if enable else 0), self._h, "set_kernel_timing")

    def take_kernel_times(self):
        """{kind: (summed ms, launches)} since the previous call (HIP events on
        the context's stream around each traversal / sort launch)."""
        ms = (ctypes.c_double * len(self.KERNEL_KINDS))()
        cnt = (ctypes.c_uint64 * len(self.KERNEL_KINDS))()
        check(lib().gr_ctx_take_kernel_times(self._h, ms, cnt), self._h, "take_kernel_times")
        return {k: (ms[i], cnt[i]) for i, k in enumerate(self.KERNEL_KINDS)}

    def close(self):
        if self._h:
            lib().gr_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Stage:
    """Scene + BVH on one device: Stage::new(prims) (src/render/stage.rs:16-21)."""

    def __init__(self, spec, ctx=None, device=0):
        self.ctx = ctx if ctx is not None else Context(device)
        self.spec = spec
        L = lib()
        self._h = ctypes.c_void_p()
        check(L.gr_scene_create(self.ctx.handle, ctypes.byref(self._h)), self.ctx.handle, "scene_create")
        for prim in spec.prims:
            mat = np.ascontiguousarray(prim.material, np.float32)
            out = ctypes.c_uint32()
            if prim.kind == "mesh":
                v = np.ascontiguousarray(prim.vertices, np.float32)
                n = None if prim.normals is None else np.ascontiguousarray(prim.normals, np.float32)
                uv = None if prim.uvs is None else np.ascontiguousarray(prim.uvs, np.float32)
                t = np.ascontiguousarray(prim.tris, np.uint32)
                code = L.gr_scene_add_mesh(self._h, _p(v, _f32p), len(v), _p(n, _f32p),
                                           0 if n is None else len(n), _p(uv, _f32p),
                                           0 if uv is None else len(uv), _p(t, _u32p), len(t),
                                           _p(mat, _f32p), ctypes.byref(out))
            else:
                c = np.ascontiguousarray(prim.center, np.float32)
                code = L.gr_scene_add_sphere(self._h, _p(c, _f32p), float(prim.radius),
                                             _p(mat, _f32p), ctypes.byref(out))
            check(code, self.ctx.handle, "scene_add")
        check(L.gr_scene_commit(self._h), self.ctx.handle, "scene_commit")

    @property
    def handle(self):
        return self._h

    def info(self):
        inf = _lib.SceneInfo()
        check(lib().gr_scene_info_get(self._h, ctypes.byref(inf)), self.ctx.handle, "scene_info")
        return {k: getattr(inf, k) for k, _ in _lib.SceneInfo._fields_}

    def export_bvh(self):
        inf = self.info()
        nn, nc = inf["num_nodes"], inf["num_components"]
        bbox = np.zeros((nn, 6), np.float32)
        off = np.zeros(nn, np.uint32)
        cnt = np.zeros(nn, np.uint32)
        axis = np.zeros(nn, np.uint32)
        cp = np.zeros(nc, np.uint32)
        ci = np.zeros(nc, np.uint32)
        check(lib().gr_scene_export_bvh(self._h, _p(bbox, _f32p), _p(off, _u32p), _p(cnt, _u32p),
                                        _p(axis, _u32p), _p(cp, _u32p), _p(ci, _u32p)),
              self.ctx.handle, "export_bvh")
        return dict(bbox=bbox, offset=off, count=cnt, axis=axis, comp_prim=cp, comp_index=ci)

    # -- batched traversal on device buffers --------------------------------
    def intersect_ptr(self, rays_ptr, n, hits_ptr, counts_ptr=None):
        check(lib().gr_intersect_closest(self._h, ctypes.c_void_p(rays_ptr), n,
                                         ctypes.c_void_p(hits_ptr),
                                         ctypes.c_void_p(counts_ptr) if counts_ptr else None),
              self.ctx.handle, "intersect_closest")

    def intersect_noalloc_ptr(self, rays_ptr, n, hits_ptr, counts_ptr=None):
        """gr_intersect_closest_noalloc: the same results with no allocation
        (scratch in a static device array; n <= GR_NOALLOC_MAX_RAYS)."""
        check(lib().gr_intersect_closest_noalloc(self._h, ctypes.c_void_p(rays_ptr), n,
                                                 ctypes.c_void_p(hits_ptr),
                                                 ctypes.c_void_p(counts_ptr) if counts_ptr else None),
              self.ctx.handle, "intersect_closest_noalloc")

    def occluded_ptr(self, segs_ptr, n, vis_ptr, mode=GR_OCCLUSION_EXACT, counts_ptr=None):
        check(lib().gr_occluded(self._h, ctypes.c_void_p(segs_ptr), n, ctypes.c_void_p(vis_ptr), mode,
                                ctypes.c_void_p(counts_ptr) if counts_ptr else None),
              self.ctx.handle, "occluded")

    def intersect(self, rays, counts=False):
        """rays: (n, 8) float32 torch tensor on the context's device.  Returns
        (hits (n, 4) int32 view of gr_hit, counts or None)."""
        import torch
        n = rays.shape[0]
        hits = torch.empty((n, 4), dtype=torch.int32, device=rays.device)
        cnt = torch.empty((n, 3), dtype=torch.int32, device=rays.device) if counts else None
        self.intersect_ptr(rays.data_ptr(), n, hits.data_ptr(), cnt.data_ptr() if counts else None)
        return hits, cnt

    def visibility(self, segs, mode=GR_OCCLUSION_EXACT, counts=False):
        import torch
        n = segs.shape[0]
        vis = torch.empty((n,), dtype=torch.int32, device=segs.device)
        cnt = torch.empty((n, 3), dtype=torch.int32, device=segs.device) if counts else None
        self.occluded_ptr(segs.data_ptr(), n, vis.data_ptr(), mode, cnt.data_ptr() if counts else None)
        return vis, cnt

    def get_tuning(self):
        t = _lib.Tuning()
        check(lib().gr_scene_get_tuning(self._h, ctypes.byref(t)), self.ctx.handle, "get_tuning")
        return {k: getattr(t, k) for k, _ in _lib.Tuning._fields_}

    def set_tuning(self, sort_bits=0, lds_stack=8, waves_per_cu=0, refill=32, xcd_ranges=1, ray_chunk=128,
                   vis_refill=40):
        t = _lib.Tuning(sort_bits, lds_stack, waves_per_cu, refill, xcd_ranges, ray_chunk, vis_refill)
        check(lib().gr_scene_set_tuning(self._h, ctypes.byref(t)), self.ctx.handle, "set_tuning")
        return self.get_tuning()

    def autotune(self, rays, segments=None, mode=GR_OCCLUSION_EXACT):
        """Times traversal settings on `rays` ((n, 8) float32 device tensor) and,
        if given, visibility `segments` ((m, 8) float32) together, and keeps the
        fastest; returns the chosen setting."""
        import torch
        n = rays.shape[0] if rays is not None else 0
        dev = rays.device if rays is not None else segments.device
        hits = torch.empty((max(n, 1), 4), dtype=torch.int32, device=dev)
        t = _lib.Tuning()
        if segments is None:
            check(lib().gr_scene_autotune(self._h, ctypes.c_void_p(rays.data_ptr() if n else 0), n,
                                          ctypes.c_void_p(hits.data_ptr()), ctypes.byref(t)),
                  self.ctx.handle, "autotune")
        else:
            m = segments.shape[0]
            vis = torch.empty((max(m, 1),), dtype=torch.int32, device=dev)
            check(lib().gr_scene_autotune_mixed(self._h, ctypes.c_void_p(rays.data_ptr() if n else 0), n,
                                                ctypes.c_void_p(hits.data_ptr()),
                                                ctypes.c_void_p(segments.data_ptr() if m else 0), m,
                                                ctypes.c_void_p(vis.data_ptr()), mode, ctypes.byref(t)),
                  self.ctx.handle, "autotune")
        return {k: getattr(t, k) for k, _ in _lib.Tuning._fields_}

    def take_overflow_count(self):
        v = ctypes.c_uint64()
        check(lib().gr_scene_take_overflow_count(self._h, ctypes.byref(v)), self.ctx.handle, "overflow")
        return v.value

    def close(self):
        if getattr(self, "_h", None):
            lib().gr_scene_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def hits_to_numpy(hits):
    """(n, 4) int32 tensor -> dict(prim, comp, t, slot) numpy arrays."""
    h = hits.cpu().numpy().view(np.uint32)
    return dict(prim=h[:, 0].copy(), comp=h[:, 1].copy(), dist=h[:, 2].view(np.float32).copy(),
                slot=h[:, 3].copy())


# ---------------------------------------------------------------------------
# Rendering: Film (src/render/film.rs) and Stage::trace (src/render/stage.rs)
# ---------------------------------------------------------------------------
INTEGRATORS = {"pt": _lib.GR_INTEGRATOR_PT, "pt_nee": _lib.GR_INTEGRATOR_PT_NEE,
               "bdpt": _lib.GR_INTEGRATOR_BDPT, "guided": _lib.GR_INTEGRATOR_GUIDED,
               "display": _lib.GR_INTEGRATOR_DISPLAY, "bdpt_mis": _lib.GR_INTEGRATOR_BDPT_MIS,
               "bdpt_lt": _lib.GR_INTEGRATOR_BDPT_LT}


def make_camera(spec):
    """gr_camera from a scenes.CameraSpec (Camera::new arguments)."""
    c = _lib.Camera()
    c.focal_length = spec.focal_length
    c.horizontal_aperture = spec.horizontal_aperture
    c.vertical_aperture = spec.vertical_aperture
    c.f_stop = spec.f_stop
    for k in range(4):
        c.rotate[k] = spec.rotate[k]
    for k in range(3):
        c.translate[k] = spec.translate[k]
    return c


class Film:
    """Accumulation film on the device: per pixel (accum rgb, weight)."""

    def __init__(self, ctx, width, height):
        self.ctx = ctx
        self.width, self.height = width, height
        self._h = ctypes.c_void_p()
        check(lib().gr_film_create(ctx.handle, width, height, ctypes.byref(self._h)), ctx.handle,
              "film_create")

    @property
    def handle(self):
        return self._h

    def clear(self):
        check(lib().gr_film_clear(self._h), self.ctx.handle, "film_clear")

    def readback(self):
        out = np.zeros((self.height * self.width, 4), np.float32)
        check(lib().gr_film_readback(self._h, _p(out, _f32p)), self.ctx.handle, "film_readback")
        return out

    def readback_begin(self):
        """Starts a host copy of the film as it will be after the work queued so
        far; later traces render meanwhile.  readback_end() returns it."""
        check(lib().gr_film_readback_begin(self._h), self.ctx.handle, "film_readback_begin")

    def readback_end(self):
        out = np.zeros((self.height * self.width, 4), np.float32)
        check(lib().gr_film_readback_end(self._h, _p(out, _f32p)), self.ctx.handle, "film_readback_end")
        return out

    def upload(self, rgbw):
        a = np.ascontiguousarray(rgbw, np.float32)
        assert a.size == self.width * self.height * 4
        check(lib().gr_film_upload(self._h, _p(a, _f32p)), self.ctx.handle, "film_upload")

    def image(self):
        px = self.readback()
        return (px[:, :3] / px[:, 3:4]).reshape(self.height, self.width, 3)

    def write_exr(self, path):
        check(lib().gr_film_write_exr(self._h, path.encode()), self.ctx.handle, "write_exr")

    def close(self):
        if getattr(self, "_h", None):
            lib().gr_film_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


GUIDE_PARAMS = 27936
GUIDE_MIX_FLOATS = 40
GUIDE_RECORD_FLOATS = 12


class Guide:
    """The learned path guide (an MLP on the matrix cores, trained online)."""

    def __init__(self, ctx, seed=1, learning_rate=5e-3, alpha=0.5, batch=16384, steps_per_pass=4,
                 train_paths=1 << 16, guided_depth=0, warmup_steps=8, lobes="consistent",
                 lobe_select="uniform", select_mix=0.25, rr_threshold=0.0):
        self.ctx = ctx
        sets = {"consistent": _lib.GR_GUIDE_LOBES_CONSISTENT, "all": _lib.GR_GUIDE_LOBES_ALL,
                "diffuse": _lib.GR_GUIDE_LOBES_DIFFUSE}
        if lobes not in sets:
            raise ValueError("guided lobe set must be one of %s" % sorted(sets))
        choices = {"uniform": _lib.GR_GUIDE_SELECT_UNIFORM, "albedo": _lib.GR_GUIDE_SELECT_ALBEDO,
                   "learned": _lib.GR_GUIDE_SELECT_LEARNED}
        if lobe_select not in choices:
            raise ValueError("lobe choice must be one of %s" % sorted(choices))
        p = _lib.GuideParams(seed, learning_rate, alpha, batch, steps_per_pass, train_paths, guided_depth,
                             warmup_steps, sets[lobes], choices[lobe_select], select_mix, rr_threshold)
        self._h = ctypes.c_void_p()
        check(lib().gr_guide_create(ctx.handle, ctypes.byref(p), ctypes.byref(self._h)), ctx.handle,
              "guide_create")

    @property
    def handle(self):
        return self._h

    def get_params(self):
        out = np.zeros(GUIDE_PARAMS, np.float32)
        check(lib().gr_guide_get_params(self._h, out.ctypes.data_as(ctypes.c_void_p), GUIDE_PARAMS),
              self.ctx.handle, "guide_get_params")
        return out

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == GUIDE_PARAMS
        check(lib().gr_guide_set_params(self._h, p.ctypes.data_as(ctypes.c_void_p), GUIDE_PARAMS),
              self.ctx.handle, "guide_set_params")

    def steps(self):
        v = ctypes.c_uint64()
        check(lib().gr_guide_steps(self._h, ctypes.byref(v)), self.ctx.handle, "guide_steps")
        return v.value

    def infer(self, raw):
        """raw: (n, 8) float32 device tensor -> (n, 40) mixtures."""
        import torch
        if raw.dim() != 2 or raw.shape[1] != 8 or raw.dtype != torch.float32 or not raw.is_contiguous():
            raise ValueError("raw inputs must be a contiguous (n, 8) float32 tensor")
        out = torch.empty((raw.shape[0], GUIDE_MIX_FLOATS), dtype=torch.float32, device=raw.device)
        check(lib().gr_guide_infer(self._h, ctypes.c_void_p(raw.data_ptr()), raw.shape[0],
                                   ctypes.c_void_p(out.data_ptr())), self.ctx.handle, "guide_infer")
        return out

    def gradient(self, records):
        import torch
        out = torch.empty((GUIDE_PARAMS,), dtype=torch.float32, device=records.device)
        check(lib().gr_guide_gradient(self._h, ctypes.c_void_p(records.data_ptr()), records.shape[0],
                                      ctypes.c_void_p(out.data_ptr())), self.ctx.handle, "guide_gradient")
        return out

    def last_records(self, cap=1 << 20):
        """The last traced batch's training records as an (n, 12) float32 device
        tensor (8 raw inputs, direction, value)."""
        import torch
        out = torch.empty((max(cap, 1), 12), dtype=torch.float32, device="cuda")
        n = ctypes.c_uint32()
        check(lib().gr_guide_last_records(self._h, ctypes.c_void_p(out.data_ptr()), cap, ctypes.byref(n)),
              self.ctx.handle, "guide_last_records")
        return out[:n.value]

    def train(self, records, steps):
        check(lib().gr_guide_train(self._h, ctypes.c_void_p(records.data_ptr()), records.shape[0], steps),
              self.ctx.handle, "guide_train")

    def close(self):
        if getattr(self, "_h", None):
            lib().gr_guide_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mixture_sample(ctx, mix, u):
    """Directions (n, 4: xyz, pdf) drawn from mixtures (n, 40) with uniforms (n, 3) (device tensors)."""
    import torch
    out = torch.empty((mix.shape[0], 4), dtype=torch.float32, device=mix.device)
    check(lib().gr_guide_mixture_sample(ctx.handle, ctypes.c_void_p(mix.data_ptr()), ctypes.c_void_p(u.data_ptr()),
                                        mix.shape[0], ctypes.c_void_p(out.data_ptr())), ctx.handle,
          "mixture_sample")
    return out


def mixture_pdf(ctx, mix, dirs):
    import torch
    out = torch.empty((mix.shape[0],), dtype=torch.float32, device=mix.device)
    check(lib().gr_guide_mixture_pdf(ctx.handle, ctypes.c_void_p(mix.data_ptr()), ctypes.c_void_p(dirs.data_ptr()),
                                     mix.shape[0], ctypes.c_void_p(out.data_ptr())), ctx.handle, "mixture_pdf")
    return out


def trace(stage, film, integrator="pt", spp=1, max_depth=0, light_depth=0, seed=1, sample_offset=0,
          batch=0, sample_radiance_ptr=None, camera=None, guide=None):
    """spp iterations of Stage::trace(&camera, &integrator, &mut film)."""
    p = _lib.TraceParams()
    p.guide = guide.handle if guide is not None else None
    p.integrator = INTEGRATORS[integrator] if isinstance(integrator, str) else integrator
    p.spp = spp
    p.max_depth = max_depth
    p.light_depth = light_depth
    p.seed = seed
    p.sample_offset = sample_offset
    p.batch = batch
    p.d_sample_radiance = sample_radiance_ptr
    cam = make_camera(camera if camera is not None else stage.spec.camera)
    check(lib().gr_trace(stage.handle, ctypes.byref(cam), ctypes.byref(p), film.handle),
          stage.ctx.handle, "trace")


JOB_ORDERS = {"as_given": _lib.GR_JOBS_AS_GIVEN, "fewest_first": _lib.GR_JOBS_FEWEST_FIRST}


def autotune_jobs(ctx, stages, rays, segments):
    """gr_jobs_autotune: the traversal setting for render_jobs over `stages`,
    timed with the batch's multi-scene traversal on `rays` and `segments`
    ((n, 8) float32 device tensors whose pad word after the origin holds the
    stage's index as uint32 bits).  Returns the setting (pass it as
    render_jobs' tuning)."""
    import torch
    n, m = rays.shape[0], segments.shape[0]
    hits = torch.empty((max(n, 1), 4), dtype=torch.int32, device=rays.device)
    vis = torch.empty((max(m, 1),), dtype=torch.int32, device=rays.device)
    arr = (ctypes.c_void_p * len(stages))(*[st.handle for st in stages])
    t = _lib.Tuning()
    check(lib().gr_jobs_autotune(ctx.handle, arr, len(stages), ctypes.c_void_p(rays.data_ptr()), n,
                                 ctypes.c_void_p(hits.data_ptr()), ctypes.c_void_p(segments.data_ptr()), m,
                                 ctypes.c_void_p(vis.data_ptr()), ctypes.byref(t)), ctx.handle, "jobs_autotune")
    return {k: getattr(t, k) for k, _ in _lib.Tuning._fields_}


def render_jobs(ctx, jobs, integrator="pt_nee", path_pool=0, steps_per_poll=0, use_graph=True, tuning=None,
                job_order="as_given", sample_ring=0):
    """Render many jobs through one shared pool of in-flight paths.  jobs: list
    of dicts with stage, film, spp, max_depth, seed and optionally
    sample_offset and camera (a CameraSpec; default: the stage's).  job_order:
    "as_given" or "fewest_first" (the order jobs take samples); sample_ring:
    finished samples held before their film gather (0 = 16 x path_pool, at
    most the batch's samples).
    Returns the per-job latency in ms (batch start to film final)."""
    n = len(jobs)
    arr = (_lib.Job * n)()
    for i, j in enumerate(jobs):
        arr[i].scene = j["stage"].handle
        arr[i].film = j["film"].handle
        arr[i].camera = make_camera(j.get("camera") or j["stage"].spec.camera)
        arr[i].spp = j["spp"]
        arr[i].max_depth = j.get("max_depth", 0)
        arr[i].seed = j.get("seed", 1)
        arr[i].sample_offset = j.get("sample_offset", 0)
    p = _lib.BatchParams()
    p.integrator = INTEGRATORS[integrator] if isinstance(integrator, str) else integrator
    p.path_pool = path_pool
    p.steps_per_poll = steps_per_poll
    p.use_graph = 1 if use_graph else 0
    p.job_order = JOB_ORDERS[job_order]
    p.sample_ring = sample_ring
    if tuning:
        t = dict(sort_bits=0, lds_stack=8, waves_per_cu=0, refill=32, xcd_ranges=1, ray_chunk=128, vis_refill=40)
        t.update(tuning)
        p.tuning = _lib.Tuning(**t)
    ms = (ctypes.c_double * n)()
    check(lib().gr_render_jobs(ctx.handle, arr, n, ctypes.byref(p), ms), ctx.handle, "render_jobs")
    return list(ms)


def material_query(stage, prim, items):
    """Device evaluation of f_world / pdf_world / light_world / sample_world
    (layout documented at gr_material_query)."""
    import torch
    x = torch.from_numpy(np.ascontiguousarray(items, np.float32).reshape(-1, 24)).cuda()
    out = torch.empty((x.shape[0], 20), dtype=torch.float32, device=x.device)
    check(lib().gr_material_query(stage.handle, prim, ctypes.c_void_p(x.data_ptr()), x.shape[0],
                                  ctypes.c_void_p(out.data_ptr())), stage.ctx.handle, "material_query")
    stage.ctx.synchronize()
    return out.cpu().numpy()
