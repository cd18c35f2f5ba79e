"""OBJ meshes (Mesh::from_obj, src/geom/mesh.rs:61-188): the library's host
loader against the OBJ oracle, bit for bit, on the reference's own cone.obj
and sphere.obj and on generated files; the oracle itself pinned by closed
forms (polygon areas for the fan triangulation, exact inverses of
translations and power-of-two scales, M M^-1 = I, unit normals)."""
import os

import numpy as np
import pytest

from gammaray_amd import scenes
from gammaray_amd._lib import GrError
from oracle import obj_oracle as OO

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = [os.path.join(HERE, "golden", "obj", n) for n in ("cone.obj", "sphere.obj")]
NONE = OO.NONE


def _translation(t):
    m = np.eye(4, dtype=np.float32)
    m[3, :3] = t
    return m


def _scale(k):
    m = np.eye(4, dtype=np.float32) * np.float32(k)
    m[3, 3] = 1.0
    return m


def _rotation_y(a):
    m = np.eye(4, dtype=np.float32)
    c, s = np.float32(np.cos(a)), np.float32(np.sin(a))
    m[0, 0], m[0, 2], m[2, 0], m[2, 2] = c, -s, s, c
    return m


# the demo's dragon transform (src/bin/gammaray.rs:16-18): Mat::scale(0.9) * Mat::translation(...)
DEMO = (_scale(0.9).astype(np.float64) @ _translation((-4.0, -5.0, -100.0)).astype(np.float64)).astype(np.float32)
XFORMS = [None, _translation((1.5, -2.0, 3.25)), _scale(2.0), DEMO,
          (_rotation_y(0.7).astype(np.float64) @ _scale(0.5).astype(np.float64)).astype(np.float32)]


def _same(a, b):
    for k in ("vertices", "normals", "uvs"):
        assert a[k].shape == b[k].shape, k
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k
    assert np.array_equal(a["tris"], b["tris"])


@pytest.mark.parametrize("path", FIXTURES)
@pytest.mark.parametrize("xi", range(len(XFORMS)))
def test_loader_matches_oracle_on_reference_assets(path, xi):
    xf = XFORMS[xi]
    got = scenes.load_obj(path, xf)
    want = OO.parse(open(path).read(), xf)
    _same(got, want)


def _polygon_area(p):
    # Newell's vector area of a planar polygon
    n = np.zeros(3)
    for i in range(len(p)):
        n += np.cross(p[i], p[(i + 1) % len(p)])
    return 0.5 * np.linalg.norm(n)


def _tri_area(v, t):
    a, b, c = v[t[0]].astype(np.float64), v[t[1]].astype(np.float64), v[t[2]].astype(np.float64)
    return 0.5 * np.linalg.norm(np.cross(b - a, c - a))


def test_fan_triangulation_keeps_polygon_area_and_corners():
    text = open(FIXTURES[0]).read()
    faces = [l.split()[1:] for l in text.splitlines() if l.startswith("f ")]
    m = OO.parse(text)
    v, tris = m["vertices"], m["tris"]
    k = 0
    for f in faces:
        idx = [int(c.split("/")[0]) - 1 for c in f]
        n = len(idx)
        fan = tris[k:k + n - 2]
        k += n - 2
        assert all(t[0] == idx[0] for t in fan)            # every triangle shares corner 0
        assert sorted(set(fan[:, :3].ravel())) == sorted(set(idx))
        area = sum(_tri_area(v, t) for t in fan)
        np.testing.assert_allclose(area, _polygon_area(v[idx].astype(np.float64)), rtol=1e-6)
    assert k == len(tris)


def test_inverse_closed_forms_and_identity():
    t = (1.5, -2.0, 3.25)
    it = OO.invert(_translation(t))
    assert np.array_equal(it, _translation([-x for x in t]))
    assert np.array_equal(OO.invert(_scale(4.0)), _scale(0.25))
    rng = np.random.default_rng(3)
    for _ in range(20):
        a = rng.normal(size=(4, 4)).astype(np.float32)
        a[:, 3] = [0, 0, 0, 1]
        a[:3, :3] += 3.0 * np.eye(3, dtype=np.float32)
        if np.linalg.det(a.astype(np.float64)) <= 0.1:
            continue
        np.testing.assert_allclose(a.astype(np.float64) @ OO.invert(a).astype(np.float64), np.eye(4), atol=2e-5)
    # a determinant that is not positive gives Mat::scale(f32::MAX)
    bad = OO.invert(_scale(-1.0))
    assert bad[0, 0] == np.finfo(np.float32).max and bad[3, 3] == 1.0


def test_normals_are_unit_and_follow_the_inverse_transpose():
    path = FIXTURES[1]
    base = OO.parse(open(path).read())
    for xf in XFORMS[1:]:
        m = OO.parse(open(path).read(), xf)
        ln = np.linalg.norm(m["normals"].astype(np.float64), axis=1)
        np.testing.assert_allclose(ln, 1.0, atol=1e-6)
        # a normal stays perpendicular to the transformed surface: n' . (M^T-transformed tangent) = 0
        want = base["normals"].astype(np.float64) @ np.linalg.inv(xf[:3, :3].astype(np.float64)).T
        want /= np.linalg.norm(want, axis=1, keepdims=True)
        np.testing.assert_allclose(m["normals"], want, atol=2e-6)


def _write(tmp_path, text, name="m.obj"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_objects_corners_and_sentinels(tmp_path):
    text = "\n".join([
        "# two objects, global indices, every corner form",
        "o first", "v 0 0 0", "v 1 0 0", "v 1 1 0", "v 0 1 0",
        "vt 0 0", "vt 1 0", "vt 1 1", "vn 0 0 2", "vn 0 0 0",
        "g side", "s 1", "usemtl m", "f 1/1/1 2/2/1 3/3/1 4/1/1",   # quad: normals, uvs
        "o second", "v 0 0 1", "v 1 0 1", "v 0 1 1  # trailing comment",
        "f 5//1 6//1 7//2",                                       # a zero normal -> geometric
        "f 5/1 6/2 7",                                            # a missing uv -> zero uv
        "l 1 2", "p 3",                                           # not triangles
        "f 5 6 7\r",
    ])
    path = _write(tmp_path, text)
    got = scenes.load_obj(path)
    want = OO.parse(text)
    _same(got, want)
    t = got["tris"]
    assert t.shape == (5, 9)
    assert t[0].tolist() == [0, 1, 2, 0, 0, 0, 0, 1, 2] and t[1].tolist() == [0, 2, 3, 0, 0, 0, 0, 2, 0]
    assert t[2].tolist() == [4, 5, 6] + [NONE] * 6
    assert t[3].tolist() == [4, 5, 6] + [NONE] * 6
    assert t[4].tolist() == [4, 5, 6] + [NONE] * 6
    assert np.array_equal(got["normals"][0], [0, 0, 1]) and np.array_equal(got["normals"][1], [0, 0, 0])


@pytest.mark.parametrize("text,what", [
    ("v 0 0 0\nv 1 0 0\nf 1 2 3\n", "line 3"),          # index past the records read so far
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf -3 -2 -1\n", "line 4"),  # relative indices are not accepted
    ("v 0 0 0\nv 1 0 0\nf 1 2\n", "line 3"),            # two corners
    ("v 0 0\n", "line 1"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\n", "no triangles"),
])
def test_loader_errors_name_the_line(tmp_path, text, what):
    with pytest.raises(GrError) as e:
        scenes.load_obj(_write(tmp_path, text))
    assert what in str(e.value)


def test_missing_file_is_an_error(tmp_path):
    with pytest.raises(GrError) as e:
        scenes.load_obj(str(tmp_path / "absent.obj"))
    assert "Couldn't open OBJ file" in str(e.value)


def test_scene_add_obj_builds_the_same_tree_as_add_mesh(host_ctx):
    """gr_scene_add_obj = gr_scene_add_mesh with the loaded arrays."""
    import ctypes
    from gammaray_amd._lib import check, lib
    from gammaray_amd.api import Stage
    xf = XFORMS[3]
    spec = scenes.SceneSpec(name="objs")
    spec.add_obj(FIXTURES[0], scenes.diffuse(), xf)
    spec.add_obj(FIXTURES[1], scenes.disney(), None)
    a = Stage(spec, ctx=host_ctx).export_bvh()
    L = lib()
    h = ctypes.c_void_p()
    check(L.gr_scene_create(host_ctx.handle, ctypes.byref(h)))
    f32 = ctypes.POINTER(ctypes.c_float)
    for path, mat, m in ((FIXTURES[0], scenes.diffuse(), xf), (FIXTURES[1], scenes.disney(), None)):
        o = ctypes.c_void_p()
        xa = None if m is None else np.ascontiguousarray(m, np.float32).reshape(16)
        check(L.gr_obj_load(path.encode(), None if xa is None else xa.ctypes.data_as(f32), ctypes.byref(o), None, 0))
        mat = np.ascontiguousarray(mat, np.float32)
        check(L.gr_scene_add_obj(h, o, mat.ctypes.data_as(f32), None))
        L.gr_obj_destroy(o)
    check(L.gr_scene_commit(h))
    from gammaray_amd._lib import SceneInfo
    inf = SceneInfo()
    check(L.gr_scene_info_get(h, ctypes.byref(inf)))
    assert inf.num_nodes == len(a["bbox"]) and inf.num_components == len(a["comp_prim"])
    bbox = np.zeros((inf.num_nodes, 6), np.float32)
    off, cnt, axis = (np.zeros(inf.num_nodes, np.uint32) for _ in range(3))
    cp, ci = np.zeros(inf.num_components, np.uint32), np.zeros(inf.num_components, np.uint32)
    u32 = ctypes.POINTER(ctypes.c_uint32)
    check(L.gr_scene_export_bvh(h, bbox.ctypes.data_as(f32), off.ctypes.data_as(u32), cnt.ctypes.data_as(u32),
                                axis.ctypes.data_as(u32), cp.ctypes.data_as(u32), ci.ctypes.data_as(u32)))
    L.gr_scene_destroy(h)
    assert np.array_equal(bbox.view(np.uint32), a["bbox"].view(np.uint32))
    assert np.array_equal(cp, a["comp_prim"]) and np.array_equal(ci, a["comp_index"])
