"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the oracle is a plain-C restatement of the reference
renderer (sdao/gammaray).  Only tests, the smoke check and the benchmark's CPU
baseline leg load it -- as the checker or as the timed CPU baseline, never as
the product path.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

NONE = 0xFFFFFFFF
_f32p = ctypes.POINTER(ctypes.c_float)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_vp = ctypes.c_void_p


def build():
    """Compile liboracle.so from the C sources in this directory."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def build_native(out_path):
    """Compile the oracle for this host's CPU (-march=native) into out_path,
    for a CPU baseline timed on the machine it runs on.  Returns True on
    success (False when no C compiler is available)."""
    try:
        r = subprocess.run(["make", "-s", "-B", "-C", _HERE, "ARCH=-march=native", "OUT=" + out_path],
                           capture_output=True, text=True, timeout=300)
    except (OSError, subprocess.TimeoutExpired):
        return False
    return r.returncode == 0 and os.path.exists(out_path)


def use_library(path):
    """Load the oracle from `path` instead of the prebuilt liboracle.so; must
    be called before the oracle is first used."""
    global _LIB_PATH
    if _lib is not None:
        raise RuntimeError("the oracle library is already loaded")
    _LIB_PATH = path


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built; run `make -C oracle`")
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_scene_create.argtypes = [ctypes.POINTER(_vp)]
        L.orc_scene_destroy.argtypes = [_vp]
        L.orc_scene_add_mesh.argtypes = [_vp, _f32p, ctypes.c_uint32, _f32p, ctypes.c_uint32,
                                         _f32p, ctypes.c_uint32, _u32p, ctypes.c_uint32, _f32p]
        L.orc_scene_add_sphere.argtypes = [_vp, _f32p, ctypes.c_float, _f32p]
        L.orc_scene_build.argtypes = [_vp]
        L.orc_bvh_num_nodes.argtypes = [_vp]
        L.orc_bvh_num_nodes.restype = ctypes.c_uint32
        L.orc_bvh_num_components.argtypes = [_vp]
        L.orc_bvh_num_components.restype = ctypes.c_uint32
        L.orc_bvh_export.argtypes = [_vp, _f32p, _u32p, _u32p, _u32p, _u32p, _u32p]
        L.orc_num_lights.argtypes = [_vp]
        L.orc_num_lights.restype = ctypes.c_uint32
        L.orc_light_indices.argtypes = [_vp, _u32p]
        L.orc_intersect.argtypes = [_vp, _f32p, ctypes.c_int64, _u32p, _u32p, _f32p, _u32p,
                                    _f32p, ctypes.c_int]
        L.orc_intersect.restype = ctypes.c_int64
        L.orc_intersect_naive.argtypes = [_vp, _f32p, ctypes.c_int64, _u32p, _u32p, _f32p,
                                          ctypes.c_int]
        L.orc_visibility.argtypes = [_vp, _f32p, ctypes.c_int64, _u8p, _u32p, ctypes.c_int, ctypes.c_int]
        L.orc_visibility.restype = ctypes.c_int64
        L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.orc_bbox_intersect.argtypes = [_f32p, _f32p, _f32p, ctypes.c_int64, _u8p]
        L.orc_camera_matrix.argtypes = [_f32p, _f32p, _f32p]
        L.orc_camera_rays.argtypes = [_f32p, ctypes.c_float, ctypes.c_float, _f32p, ctypes.c_int64, _f32p]
        L.orc_mitchell1.argtypes = [ctypes.c_float]
        L.orc_mitchell1.restype = ctypes.c_float
        L.orc_sample_points.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, _f32p]
        L.orc_expected_weights.argtypes = [ctypes.c_uint32, _f32p]
        L.orc_report_samples.argtypes = [_f32p, ctypes.c_uint32, ctypes.c_uint32, _f32p, _f32p,
                                         ctypes.c_int64]
        L.orc_render.argtypes = [_vp, _f32p, ctypes.c_float, ctypes.c_float, ctypes.c_uint32,
                                 ctypes.c_uint32, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                                 ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, _f32p, _f32p,
                                 ctypes.c_int]
        L.orc_material_query.argtypes = [_vp, ctypes.c_uint32, ctypes.c_int64, _f32p, _f32p]
        L.orc_exr_encode.argtypes = [_f32p, ctypes.c_uint32, ctypes.c_uint32, _u8p, ctypes.c_size_t]
        L.orc_exr_encode.restype = ctypes.c_size_t
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def philox4x32_10(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_p(c, _u32p), _p(k, _u32p), _p(out, _u32p))
    return out


def bbox_intersect(boxes, rays, max_dist):
    """BBox::intersect (src/core/bbox.rs:78-113) on (n, 6) boxes (lo, hi), (n, 6)
    rays (o, d) and n max distances; returns n bools."""
    b = np.ascontiguousarray(boxes, dtype=np.float32)
    r = np.ascontiguousarray(rays, dtype=np.float32)
    m = np.ascontiguousarray(max_dist, dtype=np.float32)
    n = len(b)
    assert b.shape == (n, 6) and r.shape == (n, 6) and m.shape == (n,)
    out = np.zeros(n, dtype=np.uint8)
    lib().orc_bbox_intersect(_p(b, _f32p), _p(r, _f32p), _p(m, _f32p), n, _p(out, _u8p))
    return out.astype(bool)


def default_threads():
    return max(1, min(16, os.cpu_count() or 1))


class OracleScene:
    """The oracle's Bvh over a scene description (gammaray_amd.scenes.SceneSpec)."""

    def __init__(self, spec, build=True):
        L = lib()
        self._h = _vp()
        if L.orc_scene_create(ctypes.byref(self._h)) != 0:
            raise MemoryError("orc_scene_create")
        self._keep = []
        for prim in spec.prims:
            mat = np.ascontiguousarray(prim.material, dtype=np.float32)
            if prim.kind == "mesh":
                v = np.ascontiguousarray(prim.vertices, dtype=np.float32)
                n = np.ascontiguousarray(prim.normals, dtype=np.float32) if prim.normals is not None else np.zeros((0, 3), np.float32)
                uv = np.ascontiguousarray(prim.uvs, dtype=np.float32) if prim.uvs is not None else np.zeros((0, 2), np.float32)
                t = np.ascontiguousarray(prim.tris, dtype=np.uint32)
                r = L.orc_scene_add_mesh(self._h, _p(v, _f32p), len(v), _p(n, _f32p), len(n),
                                         _p(uv, _f32p), len(uv), _p(t, _u32p), len(t), _p(mat, _f32p))
            else:
                c = np.ascontiguousarray(prim.center, dtype=np.float32)
                r = L.orc_scene_add_sphere(self._h, _p(c, _f32p), float(prim.radius), _p(mat, _f32p))
            if r < 0:
                raise ValueError("oracle rejected prim (%d)" % r)
        if build:
            if L.orc_scene_build(self._h) != 0:
                raise ValueError("orc_scene_build failed")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.orc_scene_destroy(h)
            self._h = None

    def export_bvh(self):
        L = lib()
        nn = L.orc_bvh_num_nodes(self._h)
        nc = L.orc_bvh_num_components(self._h)
        bbox = np.zeros((nn, 6), np.float32)
        off = np.zeros(nn, np.uint32)
        cnt = np.zeros(nn, np.uint32)
        axis = np.zeros(nn, np.uint32)
        cp = np.zeros(nc, np.uint32)
        ci = np.zeros(nc, np.uint32)
        L.orc_bvh_export(self._h, _p(bbox, _f32p), _p(off, _u32p), _p(cnt, _u32p), _p(axis, _u32p),
                         _p(cp, _u32p), _p(ci, _u32p))
        return dict(bbox=bbox, offset=off, count=cnt, axis=axis, comp_prim=cp, comp_index=ci)

    def light_indices(self):
        L = lib()
        n = L.orc_num_lights(self._h)
        out = np.zeros(n, np.uint32)
        if n:
            L.orc_light_indices(self._h, _p(out, _u32p))
        return out

    def intersect(self, rays, counts=False, frame=False, threads=None):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        n = len(rays)
        prim = np.zeros(n, np.uint32)
        comp = np.zeros(n, np.uint32)
        dist = np.zeros(n, np.float32)
        cnt = np.zeros((n, 3), np.uint32) if counts else None
        fr = np.zeros((n, 12), np.float32) if frame else None
        ov = lib().orc_intersect(self._h, _p(rays, _f32p), n, _p(prim, _u32p), _p(comp, _u32p),
                                 _p(dist, _f32p), _p(cnt, _u32p), _p(fr, _f32p),
                                 threads or default_threads())
        res = dict(prim=prim, comp=comp, dist=dist, overflow=int(ov))
        if counts:
            res["counts"] = cnt
        if frame:
            res["frame"] = fr
        return res

    def intersect_naive(self, rays, threads=None):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        n = len(rays)
        prim = np.zeros(n, np.uint32)
        comp = np.zeros(n, np.uint32)
        dist = np.zeros(n, np.float32)
        lib().orc_intersect_naive(self._h, _p(rays, _f32p), n, _p(prim, _u32p), _p(comp, _u32p),
                                  _p(dist, _f32p), threads or default_threads())
        return dict(prim=prim, comp=comp, dist=dist)

    def visibility(self, segs, counts=False, threads=None, early_exit=False):
        """Bvh::visibility per segment.  early_exit=True stops each traversal at
        its first occluder (the same answers; `counts` then measure that
        shorter traversal)."""
        segs = np.ascontiguousarray(segs, dtype=np.float32).reshape(-1, 6)
        n = len(segs)
        out = np.zeros(n, np.uint8)
        cnt = np.zeros((n, 3), np.uint32) if counts else None
        ov = lib().orc_visibility(self._h, _p(segs, _f32p), n, _p(out, _u8p), _p(cnt, _u32p),
                                  1 if early_exit else 0, threads or default_threads())
        res = dict(visible=out, overflow=int(ov))
        if counts:
            res["counts"] = cnt
        return res


# ---- rendering -------------------------------------------------------------
PT, PT_NEE, BDPT, DISPLAY, BDPT_MIS, BDPT_LT = 0, 1, 2, 4, 5, 6


def camera_matrix(cam):
    """Camera::new's xform (row-major 4x4) for a gammaray_amd.scenes.CameraSpec."""
    q = np.ascontiguousarray(cam.rotate, np.float32)
    t = np.ascontiguousarray(cam.translate, np.float32)
    out = np.zeros(16, np.float32)
    lib().orc_camera_matrix(_p(q, _f32p), _p(t, _f32p), _p(out, _f32p))
    return out


def window_max(cam):
    """Camera::window_max, src/core/camera.rs:76-79 (f32)."""
    f = np.float32(cam.focal_length)
    return (np.float32(cam.horizontal_aperture) / (f * np.float32(2.0)),
            np.float32(cam.vertical_aperture) / (f * np.float32(2.0)))


def camera_rays(cam, st):
    st = np.ascontiguousarray(st, np.float32).reshape(-1, 2)
    m = camera_matrix(cam)
    wx, wy = window_max(cam)
    out = np.zeros((len(st), 6), np.float32)
    lib().orc_camera_rays(_p(m, _f32p), float(wx), float(wy), _p(st, _f32p), len(st), _p(out, _f32p))
    return out


def mitchell1(x):
    return lib().orc_mitchell1(float(x))


def sample_points(seed, iteration, width, height):
    st = np.zeros((height * width, 2), np.float32)
    lib().orc_sample_points(seed, iteration, width, height, _p(st, _f32p))
    return st


def expected_weights(n):
    """Expected per-iteration film weight of each of n pixel columns (or rows)."""
    out = np.zeros(n, np.float32)
    lib().orc_expected_weights(n, _p(out, _f32p))
    return out


def report_samples(film, width, height, st, colors):
    film = np.ascontiguousarray(film, np.float32)
    st = np.ascontiguousarray(st, np.float32).reshape(-1, 2)
    colors = np.ascontiguousarray(colors, np.float32).reshape(-1, 3)
    assert len(st) == len(colors)
    lib().orc_report_samples(_p(film, _f32p), width, height, _p(st, _f32p), _p(colors, _f32p), len(st))
    return film


def exr_encode(rgbw, width, height):
    rgbw = np.ascontiguousarray(rgbw, np.float32)
    n = lib().orc_exr_encode(_p(rgbw, _f32p), width, height, None, 0)
    out = np.zeros(n, np.uint8)
    lib().orc_exr_encode(_p(rgbw, _f32p), width, height, _p(out, _u8p), n)
    return out.tobytes()


def render(scene, spec, width, height, integrator, spp, max_depth=0, light_depth=0, seed=1,
           sample_offset=0, film=None, colors=False, threads=None):
    """spp iterations of Stage::trace on the oracle.  Returns (film (H*W, 4),
    per-sample radiance (spp, H*W, 3) or None)."""
    m = camera_matrix(spec.camera)
    wx, wy = window_max(spec.camera)
    f = np.zeros((height * width, 4), np.float32) if film is None else np.ascontiguousarray(film, np.float32)
    c = np.zeros((spp, height * width, 3), np.float32) if colors else None
    r = lib().orc_render(scene._h, _p(m, _f32p), float(wx), float(wy), width, height, integrator,
                         max_depth, light_depth, seed, sample_offset, spp, _p(f, _f32p), _p(c, _f32p),
                         threads or default_threads())
    if r != 0:
        raise RuntimeError("orc_render failed")
    return f, c


def material_query(scene, prim, items):
    items = np.ascontiguousarray(items, np.float32).reshape(-1, 24)
    out = np.zeros((len(items), 20), np.float32)
    lib().orc_material_query(scene._h, prim, len(items), _p(items, _f32p), _p(out, _f32p))
    return out
