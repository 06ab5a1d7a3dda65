"""Line-by-line restatement of the reference's OBJ / OBJN reader and writer.

TEST INFRASTRUCTURE ONLY: the checker of the library's parallel reader and of
its writer (tests/test_obj.py).  Nothing in the library imports it.

read(path) follows FileObj<N> (src/obj/file/file_load.cpp):
  obj_file_dimension_and_type   src/obj/file/obj_file.cpp:144-199
  split (keyword / rest / '#')  file_load.cpp:617-686
  string_to_floats (strtof)     file_load.cpp:448-498 (libc strtof via ctypes)
  read_float / _texture         file_load.cpp:517-544
  read_integer                  file_load.cpp:272-296
  read_digit_groups             file_load.cpp:300-373
  check_index_consistent        file_load.cpp:377-398
  read_facets (fan-out)         file_load.cpp:400-446
  normalize (float, fma)        src/com/vec.h:259-290
  correct_indices               file_load.cpp:1046-1065
  check_facet_indices           file_load.cpp:851-897
  center_and_length             src/obj/alg/alg.h:36-70, 229-258
  remove_facets_with_incorrect_dimension  file_load.cpp:688-716, 902-940
write(path, ...) follows save_obj_geometry_to_file (src/obj/file/file_save.cpp:36-306).
Errors raise ObjError with the reference's message (reader line errors are
prefixed "Line k: ", k counted from 0).
"""
import ctypes
import math

import numpy as np

_libc = ctypes.CDLL(None, use_errno=True)
_libc.strtof.restype = ctypes.c_float
_libc.strtof.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p)]
_libm = ctypes.CDLL("libm.so.6")
_libm.fmaf.restype = ctypes.c_float
_libm.fmaf.argtypes = [ctypes.c_float] * 3
_libm.fma.restype = ctypes.c_double
_libm.fma.argtypes = [ctypes.c_double] * 3
ERANGE = 34

SPACE = b" \t\n\r\v\f"


class ObjError(ValueError):
    pass


def dimension(path):
    base = path.rsplit("/", 1)[-1]
    e = base.rsplit(".", 1)[1] if "." in base else ""
    if not e:
        raise ObjError("No OBJ file extension found")
    if e == "obj":
        return 3
    if not e.startswith("obj"):
        raise ObjError("Unsupported OBJ file format " + e)
    s = e[3:]
    if not s or not s.isdigit() or not s.isascii():
        raise ObjError(f'Wrong dimension number string "{s}"')
    d = int(s)
    if d < 3:
        raise ObjError(f"Wrong dimension number {d}")
    return d


def string_to_floats(text, count):
    """strtof in sequence; stops at a missing, out-of-range or non-finite value."""
    buf = ctypes.create_string_buffer(text)
    base = ctypes.addressof(buf)
    pos = 0
    out = []
    ctypes.set_errno(0)
    for _ in range(count):
        end = ctypes.c_char_p()
        v = _libc.strtof(ctypes.c_char_p(base + pos), ctypes.byref(end))
        endpos = ctypes.cast(end, ctypes.c_void_p).value - base
        if endpos == pos or ctypes.get_errno() == ERANGE or not math.isfinite(v):
            break
        out.append(np.float32(v))
        pos = endpos
    return out


def read_integer(s, i):
    b = i
    if b < len(s) and s[b:b + 1] == b"-":
        b += 1
    e = b
    while e < len(s) and s[e:e + 1].isdigit():
        e += 1
    if e == b:
        return None, i
    return int(s[i:e]), e


def read_facets(s, N):
    max_groups = (5 if N == 3 else 1) + N - 1
    groups = []
    i = 0
    while True:
        while i < len(s) and s[i] in SPACE:
            i += 1
        if i == len(s):
            break
        if len(groups) >= max_groups:
            raise ObjError(f"Found too many facet vertices {len(groups) + 1} (max supported = {max_groups})")
        v, i = read_integer(s, i)
        if v is None:
            raise ObjError("Error read facet vertex first number")
        if v == 0:
            raise ObjError("Zero facet index")
        idx = [v, 0, 0]
        for a in (1, 2):
            if i == len(s) or s[i] in SPACE:
                continue
            if s[i:i + 1] != b"/":
                raise ObjError("Error read facet vertex number")
            i += 1
            if i == len(s) or s[i] in SPACE:
                continue
            x, i = read_integer(s, i)
            if x is not None:
                if x == 0:
                    raise ObjError("Zero facet index")
                idx[a] = x
        groups.append(idx)
    if len(groups) < N:
        raise ObjError(f"Error facet vertex count {len(groups)} (min = {N})")
    tex = sum(g[1] != 0 for g in groups)
    nrm = sum(g[2] != 0 for g in groups)
    if not (tex == 0 or tex == len(groups)):
        raise ObjError("Inconsistent facet texture indices")
    if not (nrm == 0 or nrm == len(groups)):
        raise ObjError("Inconsistent facet normal indices")
    return [[groups[0]] + groups[i + 1:i + N] for i in range(len(groups) - (N - 1))]


def normalize_f32(v):
    """Vector<N,float> normalize: r = v0*v0, r = fma(vi, vi, r), v / sqrt(r), all in float."""
    v = [np.float32(x) for x in v]
    r = np.float32(v[0] * v[0])
    for x in v[1:]:
        r = np.float32(_libm.fmaf(float(x), float(x), float(r)))
    ln = np.float32(math.sqrt(float(r)))  # the double square root rounds to the float one
    return [np.float32(x / ln) for x in v]


def _extent(vertices, facets, N):
    mn = np.full(N, np.finfo(np.float32).max, np.float32)
    mx = np.full(N, np.finfo(np.float32).min, np.float32)
    for f in facets:
        for v in f:
            mn = np.minimum(mn, vertices[v])
            mx = np.maximum(mx, vertices[v])
    if np.any(mn >= mx):
        raise ObjError("Object size error")


def _has_area(a, b, c):
    e0 = [float(np.float32(b[i] - a[i])) for i in range(3)]
    e1 = [float(np.float32(c[i] - a[i])) for i in range(3)]
    return (e0[1] * e1[2] - e0[2] * e1[1] != 0 or e0[0] * e1[2] - e0[2] * e1[0] != 0
            or e0[0] * e1[1] - e0[1] * e1[0] != 0)


def read(path):
    """Returns dict(n, vertices [V,N] f32, facets [F,N] i32, normals [VN,N] f32,
    facet_normals [F,N] i32 (-1: none), texcoords count)."""
    N = dimension(path)
    data = open(path, "rb").read()
    lines = data.split(b"\n")
    if data.endswith(b"\n"):
        lines = lines[:-1]
    vertices, normals, facets = [], [], []
    vt = 0
    for num, line in enumerate(lines):
        body = line.split(b"#", 1)[0]
        parts = body.lstrip(SPACE)
        k = 0
        while k < len(parts) and parts[k] not in SPACE:
            k += 1
        kw, rest = parts[:k], parts[k:]
        try:
            if kw in (b"v", b"vn"):
                x = string_to_floats(rest, N)
                if len(x) != N:
                    raise ObjError(f"Error read {N} floating points of float type")
                if kw == b"v":
                    vertices.append(x)
                else:
                    normals.append(normalize_f32(x))
            elif kw == b"vt":
                x = string_to_floats(rest, N)
                if len(x) not in (N - 1, N):
                    raise ObjError(f"Error read {N - 1} or {N} floating points of float type")
                if len(x) == N and x[-1] != 0:
                    raise ObjError(f"{N}-dimensional textures are not supported")
                vt += 1
            elif kw == b"f":
                for fac in read_facets(rest, N):
                    vv, tt, nn = [], [], []
                    for v, t, n in fac:
                        vv.append(v - 1 if v > 0 else len(vertices) + v)
                        tt.append(t - 1 if t > 0 else (vt + t if t < 0 else -1))
                        nn.append(n - 1 if n > 0 else (len(normals) + n if n < 0 else -1))
                    facets.append((vv, nn, tt, fac[0][1] != 0, fac[0][2] != 0))
        except ObjError as e:
            raise ObjError(f"Line {num}: {e}") from None
    if not facets:
        raise ObjError("No facets found in OBJ file")
    V, VT, VN = len(vertices), vt, len(normals)
    for vv, nn, tt, has_t, has_n in facets:
        for i in range(N):
            if vv[i] < 0 or vv[i] >= V:
                raise ObjError(f"Vertex index {vv[i]} is out of bounds [0, {V})")
            if has_t and (tt[i] < 0 or tt[i] >= VT):
                raise ObjError(f"Texture coordinate index {tt[i]} is out of bounds [0, {VT})")
            if has_n and (nn[i] < 0 or nn[i] >= VN):
                raise ObjError(f"Normal index {nn[i]} is out of bounds [0, {VN})")
    vtx = np.array(vertices, np.float32).reshape(-1, N)
    fac = [f[0] for f in facets]
    fn = [f[1] for f in facets]
    _extent(vtx, fac, N)
    if N == 3:
        keep = [i for i, f in enumerate(fac) if _has_area(vtx[f[0]], vtx[f[1]], vtx[f[2]])]
        if len(keep) != len(fac):
            if not keep:
                raise ObjError("No 2-facets found in OBJ-3 file")
            fac = [fac[i] for i in keep]
            fn = [fn[i] for i in keep]
            _extent(vtx, fac, N)
    return dict(n=N, vertices=vtx, facets=np.array(fac, np.int32).reshape(-1, N),
                normals=np.array(normals, np.float32).reshape(-1, N),
                facet_normals=np.array(fn, np.int32).reshape(-1, N), texcoords=VT)


def write(path, vertices, facets, normals=None, facet_normals=None, comment=None):
    """The text save_obj_geometry_to_file writes (returned and written to path)."""
    vertices = np.asarray(vertices, np.float32)
    facets = np.asarray(facets, np.int64)
    n = vertices.shape[1]
    base = path.rsplit("/", 1)[-1]
    e = base.rsplit(".", 1)[1] if "." in base else ""
    if not e:
        raise ObjError("No OBJ file extension found")
    if e not in (("obj" if n == 3 else f"obj{n}"), f"obj{n}"):
        raise ObjError(f"Wrong OBJ-{n} file name extension: {e}")
    with_n = normals is not None and facet_normals is not None and len(normals) > 0
    has_n = []
    for f in range(len(facets)):
        k = int(np.sum(np.asarray(facet_normals[f]) >= 0)) if with_n else 0
        if k not in (0, n):
            raise ObjError("Inconsistent facet normal indices")
        has_n.append(k == n)
    used = np.unique(facets)
    if len(used) < n:
        raise ObjError(f"Facet unique indices count {len(used)} is less than {n}")
    mn = vertices[used].min(axis=0)
    mx = vertices[used].max(axis=0)
    delta = (mx - mn).astype(np.float32)
    max_delta = np.float32(delta.max())
    if max_delta == 0:
        raise ObjError("Object vertices are equal to each other")
    scale = np.float32(np.float32(2) / max_delta)
    center = (mn + np.float32(0.5) * delta).astype(np.float32)
    out = []
    if comment:
        out.append("# " + comment.replace("\n", "\n# ") + "\n")
    for v in vertices:
        x = ((v - center).astype(np.float32) * scale).astype(np.float32)
        out.append("v" + "".join(" %12.9f" % float(c) for c in x) + "\n")
    if with_n:
        for q in np.asarray(normals, np.float32):
            u = [float(c) for c in q]
            r = u[0] * u[0]
            for c in u[1:]:
                r = _libm.fma(c, c, r)
            ln = math.sqrt(r)
            if ln == 0:
                raise ObjError("Object zero length normal")
            out.append("vn" + "".join(" %12.9f" % float(np.float32(c / ln)) for c in u) + "\n")
    for f in range(len(facets)):
        fv = [int(x) for x in facets[f]]
        if has_n[f]:
            fnv = [int(x) for x in facet_normals[f]]
            if n == 3:
                p = [[float(c) for c in vertices[i]] for i in fv]
                a = [p[1][i] - p[0][i] for i in range(3)]
                b = [p[2][i] - p[0][i] for i in range(3)]
                c = [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]
                def dot(q):
                    r = float(q[0]) * c[0]
                    r = _libm.fma(float(q[1]), c[1], r)
                    return _libm.fma(float(q[2]), c[2], r)
                if all(dot(normals[fnv[k]]) < 0 for k in range(3)):
                    fv[1], fv[2] = fv[2], fv[1]
                    fnv[1], fnv[2] = fnv[2], fnv[1]
            out.append("f" + "".join(f" {a + 1}//{b + 1}" for a, b in zip(fv, fnv)) + "\n")
        else:
            out.append("f" + "".join(f" {a + 1}" for a in fv) + "\n")
    text = "".join(out)
    with open(path, "w") as fh:
        fh.write(text)
    return text
