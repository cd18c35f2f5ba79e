"""OBJ oracle (test infrastructure only): a pure-Python restatement of
Mesh::from_obj's parse and transform (src/geom/mesh.rs:61-188) for the
tests to check the library's loader against.

  * numbers: text -> f64 -> f32, as the reference casts its parser's f64
    values (mesh.rs:102,110,118);
  * vertices: Xform::transform (src/core/xform.rs:41-47), a row vector times
    the row-major Mat, divided by w, in f32;
  * normals: Xform::transform_normal with Mat::inverted
    (src/core/matrix.rs:154-290), normalised (src/core/vector.rs:66-73),
    except nearly-zero ones (math.rs:41-43 on the squared length), kept as read;
  * faces: 1-based global indices; n corners -> the fan (0, k, k+1);
  * per triangle: normal indices only if all three corners carry a normal
    and none is nearly zero, else NONE; uv indices only if all three carry
    one, else NONE.

The n-gon triangulation of the reference's parser (the wavefront_obj 5.1.0
crate, absent here) is taken to be this fan: parity with it is unpinned; the
fan is pinned by polygon areas in the tests.  All float arithmetic is numpy
float32 in the reference's operation order.
"""
import numpy as np

NONE = 0xFFFFFFFF
EPS = np.float32(np.finfo(np.float32).eps)
f32 = np.float32


def invert(m):
    """Mat::inverted (src/core/matrix.rs:154-290) in f32."""
    m = [[f32(v) for v in row] for row in np.asarray(m, np.float32)]
    x = m
    y01 = x[0][0] * x[1][1] - x[1][0] * x[0][1]
    y02 = x[0][0] * x[2][1] - x[2][0] * x[0][1]
    y03 = x[0][0] * x[3][1] - x[3][0] * x[0][1]
    y12 = x[1][0] * x[2][1] - x[2][0] * x[1][1]
    y13 = x[1][0] * x[3][1] - x[3][0] * x[1][1]
    y23 = x[2][0] * x[3][1] - x[3][0] * x[2][1]
    z33 = x[0][2] * y12 - x[1][2] * y02 + x[2][2] * y01
    z23 = x[1][2] * y03 - x[3][2] * y01 - x[0][2] * y13
    z13 = x[0][2] * y23 - x[2][2] * y03 + x[3][2] * y02
    z03 = x[2][2] * y13 - x[3][2] * y12 - x[1][2] * y23
    z32 = x[1][3] * y02 - x[2][3] * y01 - x[0][3] * y12
    z22 = x[0][3] * y13 - x[1][3] * y03 + x[3][3] * y01
    z12 = x[2][3] * y03 - x[3][3] * y02 - x[0][3] * y23
    z02 = x[1][3] * y23 - x[2][3] * y13 + x[3][3] * y12
    y01 = x[0][2] * x[1][3] - x[1][2] * x[0][3]
    y02 = x[0][2] * x[2][3] - x[2][2] * x[0][3]
    y03 = x[0][2] * x[3][3] - x[3][2] * x[0][3]
    y12 = x[1][2] * x[2][3] - x[2][2] * x[1][3]
    y13 = x[1][2] * x[3][3] - x[3][2] * x[1][3]
    y23 = x[2][2] * x[3][3] - x[3][2] * x[2][3]
    z30 = x[1][1] * y02 - x[2][1] * y01 - x[0][1] * y12
    z20 = x[0][1] * y13 - x[1][1] * y03 + x[3][1] * y01
    z10 = x[2][1] * y03 - x[3][1] * y02 - x[0][1] * y23
    z00 = x[1][1] * y23 - x[2][1] * y13 + x[3][1] * y12
    z31 = x[0][0] * y12 - x[1][0] * y02 + x[2][0] * y01
    z21 = x[1][0] * y03 - x[3][0] * y01 - x[0][0] * y13
    z11 = x[0][0] * y23 - x[2][0] * y03 + x[3][0] * y02
    z01 = x[2][0] * y13 - x[3][0] * y12 - x[1][0] * y23
    det = x[3][0] * z30 + x[2][0] * z20 + x[1][0] * z10 + x[0][0] * z00
    if not det > EPS:
        out = np.zeros((4, 4), np.float32)
        out[0, 0] = out[1, 1] = out[2, 2] = np.finfo(np.float32).max
        out[3, 3] = 1.0
        return out
    r = f32(1.0) / det
    cof = [[z00, z10, z20, z30], [z01, z11, z21, z31], [z02, z12, z22, z32], [z03, z13, z23, z33]]
    return np.array([[c * r for c in row] for row in cof], np.float32)


def _nearly_zero_vec(v):
    return abs(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) < EPS


def parse(text, xform=None):
    """dict(vertices, normals, uvs, tris) as the library's loader returns them."""
    m = np.eye(4, dtype=np.float32) if xform is None else np.asarray(xform, np.float32).reshape(4, 4)
    mm = [[f32(v) for v in row] for row in m]
    inv = invert(m)
    iv = [[f32(v) for v in row] for row in inv]
    verts, normals, uvs, tris = [], [], [], []
    for lineno, raw in enumerate(text.splitlines(), 1):
        tok = raw.split("#", 1)[0].split()
        if not tok:
            continue
        key, args = tok[0], tok[1:]
        if key == "v":
            x, y, z = (f32(float(a)) for a in args[:3])
            tx = x * mm[0][0] + y * mm[1][0] + z * mm[2][0] + mm[3][0]
            ty = x * mm[0][1] + y * mm[1][1] + z * mm[2][1] + mm[3][1]
            tz = x * mm[0][2] + y * mm[1][2] + z * mm[2][2] + mm[3][2]
            tw = x * mm[0][3] + y * mm[1][3] + z * mm[2][3] + mm[3][3]
            verts.append((tx / tw, ty / tw, tz / tw))
        elif key == "vt":
            uvs.append((f32(float(args[0])), f32(float(args[1])) if len(args) > 1 else f32(0.0)))
        elif key == "vn":
            n = tuple(f32(float(a)) for a in args[:3])
            if _nearly_zero_vec(n):
                normals.append(n)
            else:
                x, y, z = n
                nx = x * iv[0][0] + y * iv[0][1] + z * iv[0][2]
                ny = x * iv[1][0] + y * iv[1][1] + z * iv[1][2]
                nz = x * iv[2][0] + y * iv[2][1] + z * iv[2][2]
                ln = np.sqrt(nx * nx + ny * ny + nz * nz, dtype=np.float32)
                normals.append((nx / ln, ny / ln, nz / ln))
        elif key == "f":
            corners = []
            for a in args:
                parts = a.split("/")
                v = int(parts[0])
                t = int(parts[1]) if len(parts) > 1 and parts[1] else 0
                n = int(parts[2]) if len(parts) > 2 and parts[2] else 0
                if not (1 <= v <= len(verts) and 0 <= t <= len(uvs) and 0 <= n <= len(normals)):
                    raise ValueError("OBJ parse error at line %d" % lineno)
                corners.append((v, t, n))
            if len(corners) < 3:
                raise ValueError("OBJ parse error at line %d" % lineno)
            for k in range(1, len(corners) - 1):
                cs = (corners[0], corners[k], corners[k + 1])
                has_t = all(c[1] > 0 for c in cs)
                has_n = all(c[2] > 0 for c in cs) and not any(_nearly_zero_vec(normals[c[2] - 1]) for c in cs)
                tris.append([c[0] - 1 for c in cs] + [c[2] - 1 if has_n else NONE for c in cs]
                            + [c[1] - 1 if has_t else NONE for c in cs])
    if not tris:
        raise ValueError("OBJ file has no triangles")
    return dict(vertices=np.array(verts, np.float32).reshape(-1, 3),
                normals=np.array(normals, np.float32).reshape(-1, 3),
                uvs=np.array(uvs, np.float32).reshape(-1, 2),
                tris=np.array(tris, np.uint32).reshape(-1, 9))
