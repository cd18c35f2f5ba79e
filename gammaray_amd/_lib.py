"""ctypes binding of libgammaray_amd.so (the C ABI in include/gammaray.h).

Loading fails loudly: there is no CPU fallback for any entry point.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgammaray_amd.so")

GR_OK = 0
GR_HOST_ONLY = -1
GR_NONE = 0xFFFFFFFF
GR_OCCLUSION_EXACT = 0
GR_OCCLUSION_ANYHIT = 1
GR_NOALLOC_MAX_RAYS = 1 << 20
ERRORS = {-1: "GR_ERR_INVALID", -2: "GR_ERR_HIP", -3: "GR_ERR_NOMEM", -4: "GR_ERR_STATE",
          -5: "GR_ERR_OVERFLOW"}

_vp = ctypes.c_void_p
_f32p = ctypes.POINTER(ctypes.c_float)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_i = ctypes.c_int
_u32 = ctypes.c_uint32


class SceneInfo(ctypes.Structure):
    _fields_ = [("num_prims", _u32), ("num_components", _u32), ("num_nodes", _u32),
                ("num_interior", _u32), ("num_lights", _u32), ("max_depth", _u32),
                ("node_bytes", ctypes.c_uint64), ("component_bytes", ctypes.c_uint64),
                ("build_seconds", ctypes.c_double)]


class Stats(ctypes.Structure):
    _fields_ = [("nonfinite_bsdf_samples", ctypes.c_uint64), ("nonfinite_radiance", ctypes.c_uint64),
                ("bad_splats", ctypes.c_uint64), ("reserved", ctypes.c_uint64)]


class Camera(ctypes.Structure):
    _fields_ = [("focal_length", ctypes.c_float), ("horizontal_aperture", ctypes.c_float),
                ("vertical_aperture", ctypes.c_float), ("f_stop", ctypes.c_float),
                ("rotate", ctypes.c_float * 4), ("translate", ctypes.c_float * 3)]


class Tuning(ctypes.Structure):
    _fields_ = [("sort_bits", _u32), ("lds_stack", _u32), ("waves_per_cu", _u32), ("refill", _u32),
                ("xcd_ranges", _u32), ("ray_chunk", _u32), ("vis_refill", _u32)]


class TraceParams(ctypes.Structure):
    _fields_ = [("integrator", _u32), ("spp", _u32), ("max_depth", _u32), ("light_depth", _u32),
                ("seed", ctypes.c_uint64), ("sample_offset", ctypes.c_uint64), ("batch", _u32),
                ("reserved", _u32), ("d_sample_radiance", _vp), ("guide", _vp)]


class Job(ctypes.Structure):
    _fields_ = [("scene", _vp), ("film", _vp), ("camera", Camera), ("spp", _u32), ("max_depth", _u32),
                ("seed", ctypes.c_uint64), ("sample_offset", ctypes.c_uint64)]


class BatchParams(ctypes.Structure):
    _fields_ = [("integrator", _u32), ("path_pool", _u32), ("steps_per_poll", _u32), ("use_graph", _u32),
                ("tuning", Tuning), ("job_order", _u32), ("sample_ring", _u32)]


GR_JOBS_AS_GIVEN, GR_JOBS_FEWEST_FIRST = 0, 1


class GuideParams(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("learning_rate", ctypes.c_float), ("alpha", ctypes.c_float),
                ("batch", _u32), ("steps_per_pass", _u32), ("train_paths", _u32), ("guided_depth", _u32),
                ("warmup_steps", _u32), ("lobes", _u32), ("lobe_select", _u32), ("select_mix", ctypes.c_float),
                ("rr_threshold", ctypes.c_float)]


GR_INTEGRATOR_PT, GR_INTEGRATOR_PT_NEE, GR_INTEGRATOR_BDPT, GR_INTEGRATOR_GUIDED = 0, 1, 2, 3
GR_INTEGRATOR_DISPLAY = 4
GR_INTEGRATOR_BDPT_MIS = 5
GR_INTEGRATOR_BDPT_LT = 6
GR_GUIDE_LOBES_CONSISTENT, GR_GUIDE_LOBES_ALL, GR_GUIDE_LOBES_DIFFUSE = 0, 1, 2
GR_GUIDE_SELECT_UNIFORM, GR_GUIDE_SELECT_ALBEDO, GR_GUIDE_SELECT_LEARNED = 0, 1, 2

# name -> (restype, argtypes); every symbol declared in include/gammaray.h.
SIGNATURES = {
    "gr_film_create": (_i, [_vp, _u32, _u32, ctypes.POINTER(_vp)]),
    "gr_film_destroy": (None, [_vp]),
    "gr_film_clear": (_i, [_vp]),
    "gr_film_readback": (_i, [_vp, _f32p]),
    "gr_film_upload": (_i, [_vp, _f32p]),
    "gr_trace": (_i, [_vp, ctypes.POINTER(Camera), ctypes.POINTER(TraceParams), _vp]),
    "gr_film_write_exr": (_i, [_vp, ctypes.c_char_p]),
    "gr_material_query": (_i, [_vp, _u32, _vp, _u32, _vp]),
    "gr_ctx_create": (_i, [_i, ctypes.POINTER(_vp)]),
    "gr_ctx_destroy": (None, [_vp]),
    "gr_ctx_set_stream": (_i, [_vp, _vp]),
    "gr_ctx_use_own_stream": (_i, [_vp]),
    "gr_ctx_stream": (_vp, [_vp]),
    "gr_ctx_synchronize": (_i, [_vp]),
    "gr_ctx_last_error": (ctypes.c_char_p, [_vp]),
    "gr_ctx_take_stats": (_i, [_vp, ctypes.POINTER(Stats)]),
    "gr_scene_create": (_i, [_vp, ctypes.POINTER(_vp)]),
    "gr_scene_destroy": (None, [_vp]),
    "gr_scene_add_mesh": (_i, [_vp, _f32p, _u32, _f32p, _u32, _f32p, _u32, _u32p, _u32, _f32p, _u32p]),
    "gr_scene_add_sphere": (_i, [_vp, _f32p, ctypes.c_float, _f32p, _u32p]),
    "gr_film_readback_begin": (_i, [_vp]),
    "gr_film_readback_end": (_i, [_vp, _f32p]),
    "gr_exr_write": (_i, [_f32p, _u32, _u32, ctypes.c_char_p]),
    "gr_obj_load": (_i, [ctypes.c_char_p, _f32p, _vp, ctypes.c_char_p, ctypes.c_size_t]),
    "gr_obj_sizes": (_i, [_vp, _u32p, _u32p, _u32p, _u32p]),
    "gr_obj_read": (_i, [_vp, _f32p, _f32p, _f32p, _u32p]),
    "gr_obj_destroy": (None, [_vp]),
    "gr_scene_add_obj": (_i, [_vp, _vp, _f32p, _u32p]),
    "gr_scene_commit": (_i, [_vp]),
    "gr_scene_info_get": (_i, [_vp, ctypes.POINTER(SceneInfo)]),
    "gr_scene_export_bvh": (_i, [_vp, _f32p, _u32p, _u32p, _u32p, _u32p, _u32p]),
    "gr_intersect_closest": (_i, [_vp, _vp, _u32, _vp, _vp]),
    "gr_intersect_closest_noalloc": (_i, [_vp, _vp, _u32, _vp, _vp]),
    "gr_occluded": (_i, [_vp, _vp, _u32, _vp, _i, _vp]),
    "gr_scene_take_overflow_count": (_i, [_vp, _u64p]),
    "gr_scene_set_tuning": (_i, [_vp, _vp]),
    "gr_render_jobs": (_i, [_vp, _vp, _u32, ctypes.POINTER(BatchParams), _vp]),
    "gr_guide_create": (_i, [_vp, ctypes.POINTER(GuideParams), _vp]),
    "gr_guide_destroy": (None, [_vp]),
    "gr_guide_get_params": (_i, [_vp, _vp, _u32]),
    "gr_guide_set_params": (_i, [_vp, _vp, _u32]),
    "gr_guide_steps": (_i, [_vp, _u64p]),
    "gr_guide_infer": (_i, [_vp, _vp, _u32, _vp]),
    "gr_guide_gradient": (_i, [_vp, _vp, _u32, _vp]),
    "gr_guide_train": (_i, [_vp, _vp, _u32, _u32]),
    "gr_guide_last_records": (_i, [_vp, _vp, _u32, _u32p]),
    "gr_guide_mixture_sample": (_i, [_vp, _vp, _vp, _u32, _vp]),
    "gr_guide_mixture_pdf": (_i, [_vp, _vp, _vp, _u32, _vp]),
    "gr_ctx_set_kernel_timing": (_i, [_vp, _i]),
    "gr_ctx_take_kernel_times": (_i, [_vp, _vp, _vp]),
    "gr_scene_get_tuning": (_i, [_vp, _vp]),
    "gr_scene_autotune": (_i, [_vp, _vp, _u32, _vp, _vp]),
    "gr_scene_autotune_mixed": (_i, [_vp, _vp, _u32, _vp, _vp, _u32, _vp, _i, _vp]),
    "gr_jobs_autotune": (_i, [_vp, _vp, _u32, _vp, _u32, _vp, _vp, _u32, _vp, _vp]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libgammaray_amd.so is not built (run `python -m gammaray_amd.build`); "
                               "there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class GrError(RuntimeError):
    pass


def check(code, ctx=None, what=""):
    if code != GR_OK:
        msg = ""
        if ctx is not None:
            raw = lib().gr_ctx_last_error(ctx)
            msg = raw.decode() if raw else ""
        raise GrError("%s failed: %s %s" % (what, ERRORS.get(code, code), msg))
