"""Seeded synthetic scenes and ray batches for the benchmark workloads.

The reference renders one hard-coded scene whose mesh asset is not shipped
(src/bin/gammaray.rs:9-19), so every workload here is generated from a seed:

* ``box_scene``       -- 256x256 plain path tracing: a room of 6 inward quads, an
                         inner block of 6 outward quads (24 triangles in all),
                         3 spheres and one quad area light of radiance 15.
* ``icosphere``       -- the incoherent ray-cast microbenchmark: an icosphere
                         (20 * 4^8 = 1,310,720 triangles at level 8) radially
                         displaced by 4-octave value noise of amplitude 0.1.
* ``ray_batch`` / ``segment_batch`` -- origins uniform in the mesh bbox scaled
                         1.2x about its centre, directions uniform on S^2.
* ``showcase_scene``  -- the 1080p scene: an exactly-100,000-triangle displaced
                         grid, 200 spheres, a ground quad and two quad lights.

Scene content is float32 numpy arrays; the same arrays feed the GPU library
and the CPU oracle, so both see identical bits.
"""
import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

MAT_FLOATS = 16
NONE = 0xFFFFFFFF


# --------------------------------------------------------------------------
# Materials: 16-float records (kind, params...), the constructor set of
# src/material/material.rs:27-59 and the Disney builder defaults of
# src/material/material.rs:276-290.
# --------------------------------------------------------------------------
def disney(base_color=(1.0, 1.0, 1.0), roughness=0.5, anisotropic=0.0, ior=1.5, metallic=0.0,
           specular_trans=0.0, specular_tint=0.0, sheen=0.0, sheen_tint=0.5, clearcoat=0.0,
           clearcoat_gloss=0.1):
    r = np.zeros(MAT_FLOATS, np.float32)
    r[0] = 0
    r[1:4] = base_color
    r[4:14] = [roughness, anisotropic, ior, metallic, specular_trans, specular_tint, sheen,
               sheen_tint, clearcoat, clearcoat_gloss]
    return r


def diffuse_light(color):
    r = np.zeros(MAT_FLOATS, np.float32)
    r[0] = 1
    r[1:4] = color
    return r


def diffuse():
    r = np.zeros(MAT_FLOATS, np.float32)
    r[0] = 2
    return r


def mirror():
    r = np.zeros(MAT_FLOATS, np.float32)
    r[0] = 3
    return r


def ior_from_specular(spec):
    """Disney 'specular' has no field in the reference builder; map it to ior
    through F0 = 0.08 * specular, ior = (1 + sqrt(F0)) / (1 - sqrt(F0))."""
    f0 = 0.08 * float(spec)
    s = math.sqrt(f0)
    return (1.0 + s) / (1.0 - s)


def random_disney(rng, full=False):
    """Per-object Disney parameters uniform in [0, 1]: base colour, metallic,
    roughness, specular (-> ior), sheen and clearcoat; with ``full`` also
    anisotropic, specular_trans, specular_tint, sheen_tint and clearcoat_gloss."""
    base = rng.random(3)
    metallic, roughness, spec, sheen, clearcoat = rng.random(5)
    kw = dict(base_color=base, metallic=metallic, roughness=roughness,
              ior=ior_from_specular(spec), sheen=sheen, clearcoat=clearcoat)
    if full:
        a, st, stint, shtint, ccg = rng.random(5)
        kw.update(anisotropic=a, specular_trans=st, specular_tint=stint, sheen_tint=shtint,
                  clearcoat_gloss=ccg)
    return disney(**kw)


# --------------------------------------------------------------------------
# Scene description
# --------------------------------------------------------------------------
@dataclass
class PrimSpec:
    kind: str                       # "mesh" | "sphere"
    material: np.ndarray
    vertices: Optional[np.ndarray] = None   # (V,3) f32
    tris: Optional[np.ndarray] = None       # (T,9) u32: a,b,c, an,bn,cn, at,bt,ct
    normals: Optional[np.ndarray] = None    # (N,3) f32
    uvs: Optional[np.ndarray] = None        # (U,2) f32
    center: Optional[np.ndarray] = None
    radius: float = 0.0

    @property
    def num_components(self):
        return len(self.tris) if self.kind == "mesh" else 1


@dataclass
class CameraSpec:
    """Camera::new arguments (src/core/camera.rs:42-61)."""
    focal_length: float = 5.0
    horizontal_aperture: float = 2.2
    vertical_aperture: float = 1.6
    f_stop: float = 8.0
    rotate: tuple = (1.0, 0.0, 0.0, 0.0)    # quaternion (real, i, j, k)
    translate: tuple = (0.0, 0.0, 0.0)


@dataclass
class SceneSpec:
    prims: List[PrimSpec] = field(default_factory=list)
    camera: CameraSpec = field(default_factory=CameraSpec)
    name: str = ""

    def add_mesh(self, vertices, faces, material, normals=None, uvs=None, face_normals=None,
                 face_uvs=None):
        faces = np.asarray(faces, dtype=np.uint32).reshape(-1, 3)
        t = np.full((len(faces), 9), NONE, dtype=np.uint32)
        t[:, 0:3] = faces
        if face_normals is not None:
            t[:, 3:6] = face_normals
        if face_uvs is not None:
            t[:, 6:9] = face_uvs
        self.prims.append(PrimSpec("mesh", np.asarray(material, np.float32),
                                   vertices=np.ascontiguousarray(vertices, np.float32), tris=t,
                                   normals=None if normals is None else np.ascontiguousarray(normals, np.float32),
                                   uvs=None if uvs is None else np.ascontiguousarray(uvs, np.float32)))
        return len(self.prims) - 1

    def add_obj(self, path, material, xform=None):
        """Mesh::from_obj(material, xform, path) (src/geom/mesh.rs:61-188): the
        file parsed by the library's OBJ loader; xform is a 4x4 row-major Mat."""
        m = load_obj(path, xform)
        self.prims.append(PrimSpec("mesh", np.asarray(material, np.float32), vertices=m["vertices"],
                                   tris=m["tris"], normals=m["normals"] if len(m["normals"]) else None,
                                   uvs=m["uvs"] if len(m["uvs"]) else None))
        return len(self.prims) - 1

    def add_sphere(self, center, radius, material):
        self.prims.append(PrimSpec("sphere", np.asarray(material, np.float32),
                                   center=np.asarray(center, np.float32), radius=float(np.float32(radius))))
        return len(self.prims) - 1

    @property
    def num_components(self):
        return sum(p.num_components for p in self.prims)

    def bbox(self):
        lo = np.full(3, np.inf)
        hi = np.full(3, -np.inf)
        for p in self.prims:
            if p.kind == "mesh":
                lo = np.minimum(lo, p.vertices.min(0))
                hi = np.maximum(hi, p.vertices.max(0))
            else:
                lo = np.minimum(lo, p.center - p.radius)
                hi = np.maximum(hi, p.center + p.radius)
        return lo, hi


def load_obj(path, xform=None):
    """An OBJ file as the arrays gr_scene_add_mesh takes: vertices (V,3),
    normals (N,3), uvs (U,2) float32 and tris (T,9) uint32 (GR_NONE = geometric
    normal / zero uv), through the library's loader (host only, no GPU)."""
    import ctypes
    from ._lib import GrError, lib
    L = lib()
    xf = None if xform is None else np.ascontiguousarray(xform, np.float32).reshape(16)
    h = ctypes.c_void_p()
    err = ctypes.create_string_buffer(512)
    code = L.gr_obj_load(str(path).encode(), None if xf is None else xf.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                         ctypes.byref(h), err, len(err))
    if code != 0:
        raise GrError("gr_obj_load failed: " + err.value.decode(errors="replace"))
    try:
        n = [ctypes.c_uint32() for _ in range(4)]
        L.gr_obj_sizes(h, *[ctypes.byref(x) for x in n])
        nv, nn, nt, ntri = (x.value for x in n)
        out = dict(vertices=np.zeros((nv, 3), np.float32), normals=np.zeros((nn, 3), np.float32),
                   uvs=np.zeros((nt, 2), np.float32), tris=np.zeros((ntri, 9), np.uint32))
        f32 = ctypes.POINTER(ctypes.c_float)
        L.gr_obj_read(h, out["vertices"].ctypes.data_as(f32), out["normals"].ctypes.data_as(f32),
                      out["uvs"].ctypes.data_as(f32), out["tris"].ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    finally:
        L.gr_obj_destroy(h)
    return out


def quad(center, u, v, normal):
    """Two triangles spanning center +- u +- v whose geometric normal
    (a - c) x (b - c) (src/geom/mesh.rs:194-204) points along ``normal``."""
    c = np.asarray(center, np.float64)
    u = np.asarray(u, np.float64)
    v = np.asarray(v, np.float64)
    if np.dot(np.cross(u, v), normal) < 0:
        u, v = v, u
    p = np.array([c - u - v, c + u - v, c + u + v, c - u + v])
    # triangle (p0, p1, p2): (p0 - p2) x (p1 - p2) is parallel to u x v
    faces = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    return p.astype(np.float32), faces


def _box_quads(lo, hi, inward):
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    c = 0.5 * (lo + hi)
    h = 0.5 * (hi - lo)
    out = []
    for axis in range(3):
        for sgn in (-1.0, 1.0):
            center = c.copy()
            center[axis] += sgn * h[axis]
            a1, a2 = [a for a in range(3) if a != axis]
            u = np.zeros(3); u[a1] = h[a1]
            v = np.zeros(3); v[a2] = h[a2]
            n = np.zeros(3); n[axis] = -sgn if inward else sgn
            out.append(quad(center, u, v, n))
    return out


# --------------------------------------------------------------------------
# 256x256 box scene (plain path tracing)
# --------------------------------------------------------------------------
def box_scene(seed=1, consistent=False, full=False):
    """consistent=True zeroes clearcoat, the one lobe of these materials whose
    reported pdf is not its sampling density (GTR1, src/material/util.rs:
    266-308); full=True draws every Disney parameter (adds transmission)."""
    rng = np.random.default_rng(seed)
    s = SceneSpec(name="box256")

    def mat():
        m = random_disney(rng, full=full)
        if consistent:
            m[12] = 0.0
        return m
    for verts, faces in _box_quads((-1.0, 0.0, -1.0), (1.0, 2.0, 1.0), inward=True):
        s.add_mesh(verts, faces, mat())
    bv, bf = [], []
    for k, (verts, faces) in enumerate(_box_quads((-0.7, 0.002, -0.7), (-0.15, 0.7, -0.2), inward=False)):
        bv.append(verts)
        bf.append(faces + 4 * k)
    s.add_mesh(np.concatenate(bv), np.concatenate(bf), mat())
    for _ in range(3):
        r = 0.15 + 0.2 * rng.random()
        c = np.array([rng.uniform(-0.9 + r, 0.9 - r), r + rng.uniform(0.0, 1.0),
                      rng.uniform(-0.9 + r, 0.2)])
        s.add_sphere(c, r, mat())
    lv, lf = quad((0.0, 1.98, 0.0), (0.3, 0.0, 0.0), (0.0, 0.0, 0.3), (0.0, -1.0, 0.0))
    s.add_mesh(lv, lf, diffuse_light((15.0, 15.0, 15.0)))
    # Square film: equal apertures; ~63 degree field of view looking down -z.
    s.camera = CameraSpec(focal_length=0.8, horizontal_aperture=1.0, vertical_aperture=1.0,
                          translate=(0.0, 1.0, 0.98))
    return s


def demo_scene(obj_path=None):
    """The reference binary's scene (src/bin/gammaray.rs:8-81), as the
    `gammaray` driver builds it: the OBJ mesh (the dragon there; skipped when
    obj_path is None) scaled by 0.9 and moved to (-4, -5, -100), a light
    sphere, a large metal sphere, a glass sphere and two rows of 20 spheres,
    seen by Camera::default."""
    s = SceneSpec(name="demo")
    if obj_path is not None:
        xf = np.eye(4, dtype=np.float32) * np.float32(0.9)
        xf[3] = (-4.0, -5.0, -100.0, 1.0)
        s.add_obj(obj_path, disney(base_color=(0.0, 0.5, 1.0), roughness=0.5, metallic=1.0, anisotropic=1.0,
                                   sheen=1.0), xf)
    s.add_sphere((12.0, 3.0, -90.0), 5.0, diffuse_light((2.0, 2.0, 2.0)))
    s.add_sphere((-25.0, 0.0, -50.0), 75.0, disney(base_color=(0.5, 0.9, 0.0), roughness=0.5, metallic=1.0))
    s.add_sphere((6.0, -10.0, -90.0), 4.0, disney(specular_trans=1.0, roughness=0.2, ior=1.8, metallic=0.0))
    for row in range(2):
        for i in range(20):
            color = (0.9 if i % 2 == 0 else 0.2, 0.9 if i % 3 == 0 else 0.2, 0.9 if i % 5 == 0 else 0.2)
            if row == 0:
                m = disney(base_color=color, roughness=0.5, metallic=1.0, anisotropic=0.8, clearcoat=1.0,
                           clearcoat_gloss=1.0)
            else:
                m = disney(base_color=color, roughness=0.2, ior=1.8, metallic=0.0, sheen=1.0)
            s.add_sphere((-20.0 + i * 4, 12.0 if row == 0 else -12.0, -100.0), 1.8, m)
    s.camera = CameraSpec()
    return s


def indirect_room_scene(seed=5):
    """A room lit only indirectly: the area light under the ceiling emits
    upwards, so every surface below it is lit by the bright patch the light
    makes on the ceiling -- light sampling finds nothing from below and BSDF
    sampling rarely hits the patch, the case a learned guide is for.
    Materials are rough Disney dielectrics without clearcoat."""
    rng = np.random.default_rng(seed)
    s = SceneSpec(name="indirect_room")
    for verts, faces in _box_quads((-1.0, 0.0, -1.0), (1.0, 1.0, 1.0), inward=True):
        s.add_mesh(verts, faces, disney(base_color=0.4 + 0.5 * rng.random(3), roughness=0.9))
    lv, lf = quad((0.4, 0.9, 0.4), (0.15, 0.0, 0.0), (0.0, 0.0, 0.15), (0.0, 1.0, 0.0))
    s.add_mesh(lv, lf, diffuse_light((40.0, 40.0, 40.0)))
    for _ in range(2):
        r = 0.12 + 0.1 * rng.random()
        s.add_sphere((rng.uniform(-0.6, 0.6), r, rng.uniform(-0.6, 0.6)), r,
                     disney(base_color=rng.random(3), roughness=0.6 + 0.4 * rng.random()))
    s.camera = CameraSpec(focal_length=0.8, horizontal_aperture=1.0, vertical_aperture=1.0,
                          translate=(0.0, 0.5, 0.98))
    return s


# --------------------------------------------------------------------------
# Displaced icosphere (ray-cast microbenchmark)
# --------------------------------------------------------------------------
def _icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def _subdivide(v, f):
    """Split each triangle into 4 with shared, normalised edge midpoints."""
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    e.sort(axis=1)
    key = e[:, 0] * (len(v) + 1) + e[:, 1]
    uniq, inv = np.unique(key, return_inverse=True)
    a = uniq // (len(v) + 1)
    b = uniq % (len(v) + 1)
    mid = v[a] + v[b]
    mid /= np.linalg.norm(mid, axis=1, keepdims=True)
    m = inv.reshape(3, -1).T + len(v)      # midpoint ids of edges (01, 12, 20)
    v2 = np.concatenate([v, mid])
    m01, m12, m20 = m[:, 0], m[:, 1], m[:, 2]
    f2 = np.concatenate([np.stack([f[:, 0], m01, m20], 1), np.stack([f[:, 1], m12, m01], 1),
                         np.stack([f[:, 2], m20, m12], 1), np.stack([m01, m12, m20], 1)])
    return v2, f2


def _hash3(ix, iy, iz, seed):
    h = (ix.astype(np.uint32) * np.uint32(0x8DA6B343) ^ iy.astype(np.uint32) * np.uint32(0xD8163841)
         ^ iz.astype(np.uint32) * np.uint32(0xCB1AB31F) ^ np.uint32(seed * 0x9E3779B9 & 0xFFFFFFFF))
    h ^= h >> np.uint32(15)
    h *= np.uint32(0x2C1B3C6D)
    h ^= h >> np.uint32(12)
    h *= np.uint32(0x297A2D39)
    h ^= h >> np.uint32(15)
    return h.astype(np.float64) / 4294967295.0 * 2.0 - 1.0


def value_noise(p, seed, octaves=4, base_freq=2.0):
    """Smoothstep-interpolated lattice value noise, octaves summed with halving
    amplitude and doubling frequency, normalised to [-1, 1]."""
    p = np.asarray(p, np.float64)
    total = np.zeros(len(p))
    amp, norm, freq = 1.0, 0.0, base_freq
    with np.errstate(over="ignore"):
        for o in range(octaves):
            q = p * freq + 17.0 * o
            i0 = np.floor(q)
            fr = q - i0
            w = fr * fr * (3.0 - 2.0 * fr)
            i0 = i0.astype(np.int64)
            acc = np.zeros(len(p))
            for dx in (0, 1):
                wx = w[:, 0] if dx else 1.0 - w[:, 0]
                for dy in (0, 1):
                    wy = w[:, 1] if dy else 1.0 - w[:, 1]
                    for dz in (0, 1):
                        wz = w[:, 2] if dz else 1.0 - w[:, 2]
                        acc += wx * wy * wz * _hash3(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz,
                                                     seed + o)
            total += amp * acc
            norm += amp
            amp *= 0.5
            freq *= 2.0
    return total / norm


def icosphere(level=8, seed=2, amplitude=0.1):
    """Displaced icosphere: (vertices f32 (V,3), faces u32 (T,3))."""
    v, f = _icosahedron()
    for _ in range(level):
        v, f = _subdivide(v, f)
    d = 1.0 + amplitude * value_noise(v, seed)
    return (v * d[:, None]).astype(np.float32), f.astype(np.uint32)


def icosphere_scene(level=8, seed=2):
    v, f = icosphere(level, seed)
    s = SceneSpec(name="icosphere_L%d" % level)
    s.add_mesh(v, f, disney(base_color=(0.8, 0.8, 0.8)))
    return s


def _uniform_dirs(rng, n):
    z = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * math.pi, n)
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)


def _box_points(rng, lo, hi, n, scale=1.2):
    c = 0.5 * (lo + hi)
    h = 0.5 * (hi - lo) * scale
    return c + rng.uniform(-1.0, 1.0, (n, 3)) * h


def ray_batch(spec, n, seed, chunk=1 << 22):
    """(n, 8) float32 rays: origin, pad, unit direction, pad -- origins uniform
    in the scene bbox scaled 1.2x, directions uniform on the sphere."""
    lo, hi = spec.bbox()
    rng = np.random.default_rng(seed)
    out = np.zeros((n, 8), np.float32)
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        out[s:s + m, 0:3] = _box_points(rng, lo, hi, m)
        d = _uniform_dirs(rng, m)
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        out[s:s + m, 4:7] = d
    return out


def segment_batch(spec, n, seed, chunk=1 << 22):
    """(n, 8) float32 shadow segments: start, pad, target, pad, both uniform in
    the scene bbox scaled 1.2x."""
    lo, hi = spec.bbox()
    rng = np.random.default_rng(seed)
    out = np.zeros((n, 8), np.float32)
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        out[s:s + m, 0:3] = _box_points(rng, lo, hi, m)
        out[s:s + m, 4:7] = _box_points(rng, lo, hi, m)
    return out


def _surface_points(spec, rng, n):
    """n points on the scene's non-emitting mesh triangles (uniform over the
    triangles, then uniform in each), with each triangle's unit normal turned
    to +y, from above which the scenes' cameras and lights see them."""
    tris = [(p.vertices[p.tris[:, 0]], p.vertices[p.tris[:, 1]], p.vertices[p.tris[:, 2]])
            for p in spec.prims if p.kind == "mesh" and p.material[0] != 1]
    a = np.concatenate([t[0] for t in tris]).astype(np.float64)
    b = np.concatenate([t[1] for t in tris]).astype(np.float64)
    c = np.concatenate([t[2] for t in tris]).astype(np.float64)
    k = rng.integers(0, len(a), n)
    u, v = rng.random(n), rng.random(n)
    flip = u + v > 1.0
    u[flip], v[flip] = 1.0 - u[flip], 1.0 - v[flip]
    pts = a[k] + (b[k] - a[k]) * u[:, None] + (c[k] - a[k]) * v[:, None]
    nrm = np.cross(b[k] - a[k], c[k] - a[k])
    nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-30)
    nrm[nrm[:, 1] < 0] *= -1.0
    return pts, nrm


def bounce_batch(spec, n, seed):
    """(n, 8) float32 rays like a path tracer's bounces off the scene's
    surfaces: origins on its mesh triangles, lifted 1e-4 along the normal,
    directions cosine-distributed about the normal."""
    rng = np.random.default_rng(seed)
    pts, nrm = _surface_points(spec, rng, n)
    r, phi = np.sqrt(rng.random(n)), 2.0 * math.pi * rng.random(n)
    local = np.stack([r * np.cos(phi), np.sqrt(np.maximum(0.0, 1.0 - r * r)), r * np.sin(phi)], 1)
    # rotate +y onto the normal (Rodrigues; the normals point up, so y + 1 > 0)
    x, y, z = nrm[:, 0:1], nrm[:, 1:2], nrm[:, 2:3]
    k = 1.0 / (1.0 + y)
    lx, ly, lz = local[:, 0:1], local[:, 1:2], local[:, 2:3]
    d = np.concatenate([lx * (y + z * z * k) + ly * x - lz * x * z * k,
                        -lx * x + ly * y - lz * z,
                        -lx * x * z * k + ly * z + lz * (y + x * x * k)], 1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    out = np.zeros((n, 8), np.float32)
    out[:, 0:3] = pts + 1e-4 * nrm
    out[:, 4:7] = d
    return out


def light_segment_batch(spec, n, seed):
    """(n, 8) float32 shadow segments like next-event estimation's: from points
    on the scene's mesh triangles to uniform points of its emitting meshes."""
    rng = np.random.default_rng(seed)
    pts, nrm = _surface_points(spec, rng, n)
    lights = [p for p in spec.prims if p.kind == "mesh" and p.material[0] == 1]
    la = np.concatenate([p.vertices[p.tris[:, 0]] for p in lights]).astype(np.float64)
    lb = np.concatenate([p.vertices[p.tris[:, 1]] for p in lights]).astype(np.float64)
    lc = np.concatenate([p.vertices[p.tris[:, 2]] for p in lights]).astype(np.float64)
    k = rng.integers(0, len(la), n)
    u, v = rng.random(n), rng.random(n)
    flip = u + v > 1.0
    u[flip], v[flip] = 1.0 - u[flip], 1.0 - v[flip]
    out = np.zeros((n, 8), np.float32)
    out[:, 0:3] = pts + 1e-4 * nrm
    out[:, 4:7] = la[k] + (lb[k] - la[k]) * u[:, None] + (lc[k] - la[k]) * v[:, None]
    return out


def rays6(r8):
    """Drop the padding lanes: (n, 8) -> (n, 6)."""
    r8 = np.asarray(r8)
    return np.ascontiguousarray(np.concatenate([r8[:, 0:3], r8[:, 4:7]], 1))


# --------------------------------------------------------------------------
# 1080p showcase scene
# --------------------------------------------------------------------------
def displaced_grid(nx, nz, size, seed, amplitude=0.25, center=(0.0, 0.0, 0.0)):
    """(nx x nz quads -> 2*nx*nz triangles) height field displaced by value noise."""
    xs = np.linspace(-0.5, 0.5, nx + 1) * size[0]
    zs = np.linspace(-0.5, 0.5, nz + 1) * size[1]
    X, Z = np.meshgrid(xs, zs, indexing="ij")
    P = np.stack([X.ravel(), np.zeros(X.size), Z.ravel()], 1)
    h = value_noise(P / max(size), seed) * amplitude * max(size)
    P[:, 1] = h
    P += np.asarray(center)
    idx = np.arange((nx + 1) * (nz + 1)).reshape(nx + 1, nz + 1)
    a = idx[:-1, :-1].ravel(); b = idx[1:, :-1].ravel(); c = idx[1:, 1:].ravel(); d = idx[:-1, 1:].ravel()
    # winding so that the geometric normal points to +y
    f = np.concatenate([np.stack([a, d, c], 1), np.stack([a, c, b], 1)]).astype(np.uint32)
    return P.astype(np.float32), f


def showcase_scene(seed=3, n_spheres=200, grid=(250, 200)):
    rng = np.random.default_rng(seed)
    s = SceneSpec(name="showcase")
    gv, gf = displaced_grid(grid[0], grid[1], (6.0, 4.8), seed, amplitude=0.06, center=(0.0, 0.4, -1.0))
    s.add_mesh(gv, gf, random_disney(rng, full=True))
    qv, qf = quad((0.0, 0.0, -1.0), (6.0, 0.0, 0.0), (0.0, 0.0, 6.0), (0.0, 1.0, 0.0))
    s.add_mesh(qv, qf, random_disney(rng, full=True))
    for _ in range(n_spheres):
        r = rng.uniform(0.05, 0.5)
        c = np.array([rng.uniform(-4.0, 4.0), rng.uniform(0.5, 3.0), rng.uniform(-5.0, 1.0)])
        s.add_sphere(c, r, random_disney(rng, full=True))
    for x in (-2.0, 2.0):
        lv, lf = quad((x, 4.5, -1.0), (0.8, 0.0, 0.0), (0.0, 0.0, 0.8), (0.0, -1.0, 0.0))
        s.add_mesh(lv, lf, diffuse_light((12.0, 11.0, 10.0)))
    s.camera = CameraSpec(focal_length=1.0, horizontal_aperture=1.6, vertical_aperture=0.9,
                          translate=(0.0, 1.6, 4.0))
    return s


# --------------------------------------------------------------------------
# Batched-rendering workload: 8 scenes, 64 jobs
# --------------------------------------------------------------------------
def terrain_scene(triangles, seed):
    """A displaced-mesh terrain of about ``triangles`` triangles (a square
    grid, 2 triangles per cell) with three spheres and a quad light above,
    Disney materials drawn per object."""
    rng = np.random.default_rng(seed)
    n = max(2, int(round(math.sqrt(triangles / 2.0))))
    s = SceneSpec(name="terrain%d" % (2 * n * n))
    gv, gf = displaced_grid(n, n, (4.0, 4.0), seed, amplitude=0.08, center=(0.0, 0.0, -2.0))
    s.add_mesh(gv, gf, random_disney(rng))
    for _ in range(3):
        r = rng.uniform(0.15, 0.4)
        s.add_sphere((rng.uniform(-1.2, 1.2), 0.5 + r, rng.uniform(-3.0, -1.0)), r, random_disney(rng))
    lv, lf = quad((0.0, 3.0, -2.0), (0.6, 0.0, 0.0), (0.0, 0.0, 0.6), (0.0, -1.0, 0.0))
    s.add_mesh(lv, lf, diffuse_light((18.0, 17.0, 16.0)))
    s.camera = CameraSpec(focal_length=1.0, horizontal_aperture=1.0, vertical_aperture=1.0,
                          rotate=(0.9659258, -0.258819, 0.0, 0.0), translate=(0.0, 1.6, 0.8))
    return s


def batch_workload(seed=4, n_jobs=64, n_scenes=8):
    """Triangle counts 10k * 200^(i/7) (10k .. 2M), and per job a side length
    log-uniform in [128, 1024], spp from {4, 16, 64, 256} and a scene, all
    uniform.  Returns (scene specs, job dicts with scene index, side, spp, seed)."""
    rng = np.random.default_rng(seed)
    specs = [terrain_scene(int(10000 * 200.0 ** (i / (n_scenes - 1))), seed * 100 + i) for i in range(n_scenes)]
    jobs = []
    for j in range(n_jobs):
        side = int(round(math.exp(rng.uniform(math.log(128), math.log(1024)))))
        jobs.append(dict(scene=int(rng.integers(0, n_scenes)), side=side,
                         spp=int(rng.choice([4, 16, 64, 256])), seed=1000 + j))
    return specs, jobs

