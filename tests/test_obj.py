"""OBJ / OBJN reading and writing (host code of the library) against the
line-by-line restatement of the reference's reader and writer (oracle/obj_ref.py).

Known-answer files pin both: their expected arrays are worked out by hand from
the format rules (fan-out of a face line, 1-based and negative indices, float
normalisation of vn, removal of degenerate triangles, the writer's scaling and
number format).  Random files then check the library's chunked parallel reader
against the oracle bit for bit, with 1, 3 and 8 chunks and with the automatic
split of a file of several MiB, and check that both raise the same message on
malformed input.  The GPU test at the end builds a mesh from a read file and
compares it with the mesh built from the same arrays.
"""
import os

import numpy as np
import pytest

from math_amd import native
from oracle import obj_ref


def lib_read(path, threads=0):
    o = native.read_obj(path, threads)
    return dict(n=o.n, vertices=o.vertices, facets=o.facets, normals=o.normals,
                facet_normals=o.facet_normals, texcoords=o.texcoords)


def assert_same(a, b):
    assert a["n"] == b["n"] and a["texcoords"] == b["texcoords"]
    for k in ("vertices", "facets", "normals", "facet_normals"):
        assert a[k].shape == b[k].shape, k
        assert a[k].tobytes() == b[k].tobytes(), k


def write_text(tmp_path, name, text):
    p = os.path.join(str(tmp_path), name)
    with open(p, "w", newline="") as fh:
        fh.write(text)
    return p


# ---------------------------------------------------------------------------
# Known answers (worked out by hand)

FAN = """# five groups on one line give three triangles around the first vertex
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 0 0 1
vn 0 0 2
vn 3 4 0
f 1 2 3 4 5
f -5//1 -4//1 -3//2
o ignored
usemtl ignored
s off
"""


def test_known_answer_fan_out_and_indices(tmp_path):
    p = write_text(tmp_path, "fan.obj", FAN)
    for r in (obj_ref.read(p), lib_read(p), lib_read(p, 3)):
        assert r["n"] == 3 and r["texcoords"] == 0
        assert r["vertices"].tolist() == [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1]]
        # group 0 with groups (1,2), (2,3), (3,4); then -5 -4 -3 of five vertices = 0 1 2
        assert r["facets"].tolist() == [[0, 1, 2], [0, 2, 3], [0, 3, 4], [0, 1, 2]]
        assert r["facet_normals"].tolist() == [[-1] * 3, [-1] * 3, [-1] * 3, [0, 0, 1]]
        # (0,0,2)/2 and (3,4,0)/5 in float
        assert r["normals"].tolist() == [[0, 0, 1], [np.float32(0.6), np.float32(0.8), 0]]


def test_known_answer_objn_and_forward_references(tmp_path):
    # 4-space: one facet per line; positive indices may name vertices read later
    text = "f 1 2 3 4\nv 0 0 0 0\nv 1 0 0 0\nv 0 1 0 0\nv 0 0 1 0\nv 0 0 0 1\nf -4 -3 -2 -1\n"
    p = write_text(tmp_path, "t.obj4", text)
    for r in (obj_ref.read(p), lib_read(p)):
        assert r["n"] == 4
        assert r["facets"].tolist() == [[0, 1, 2, 3], [1, 2, 3, 4]]


def test_known_answer_degenerate_triangles_removed(tmp_path):
    # (1 2 2) and the collinear (1 2 5) are dropped; vt with a zero third value is allowed
    text = "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 0 0 1\nv 2 0 0\nvt 0.5 0.5 0\nf 1 2 2\nf 1 2 5\nf 1 3 4\nf 2 3 4\n"
    p = write_text(tmp_path, "d.obj", text)
    for r in (obj_ref.read(p), lib_read(p)):
        assert r["facets"].tolist() == [[0, 2, 3], [1, 2, 3]]
        assert r["texcoords"] == 1


def test_known_answer_writer(tmp_path):
    v = np.array([[0, 0, 0], [4, 0, 0], [0, 2, 0], [0, 0, 1], [9, 9, 9]], np.float32)
    f = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    nrm = np.array([[0, 0, -2]], np.float32)
    fn = np.array([[0, 0, 0], [-1, -1, -1]], np.int32)
    # box of the used vertices [0,4]x[0,2]x[0,1]: centre (2,1,0.5), scale 2/4; vertex 4 is unused
    # but written; facet 0's normal points against (v1-v0)x(v2-v0) = +z, so vertices 1 and 2 swap
    expect = ("# a\n# b\n"
              "v -1.000000000 -0.500000000 -0.250000000\n"
              "v  1.000000000 -0.500000000 -0.250000000\n"
              "v -1.000000000  0.500000000 -0.250000000\n"
              "v -1.000000000 -0.500000000  0.250000000\n"
              "v  3.500000000  4.000000000  4.250000000\n"
              "vn  0.000000000  0.000000000 -1.000000000\n"
              "f 1//1 3//1 2//1\n"
              "f 1 3 4\n")
    p1 = os.path.join(str(tmp_path), "w1.obj")
    p2 = os.path.join(str(tmp_path), "w2.obj")
    assert obj_ref.write(p1, v, f, nrm, fn, comment="a\nb") == expect
    native.write_obj(p2, v, f, nrm, fn, comment="a\nb")
    assert open(p2).read() == expect


ERRORS = [
    ("e.obj", "v 0 0 0\nf 1 0 1\n", "Line 1: Zero facet index"),
    ("e.obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1//1 2 3\n", "Line 4: Inconsistent facet normal indices"),
    ("e.obj", "v 0 0 0\nv 1 0 0\nf 1 2\n", "Line 2: Error facet vertex count 2 (min = 3)"),
    ("e.obj4", "v 0 0 0 0\nf 1 2 3 4 1\n", "Line 1: Found too many facet vertices 5 (max supported = 4)"),
    ("e.obj", "v 0 0 0\nf 1 2 3 4 5 6 7 8\n", "Line 1: Found too many facet vertices 8 (max supported = 7)"),
    ("e.obj", "v 0 0\n", "Line 0: Error read 3 floating points of float type"),
    ("e.obj", "v 0 0 1e39\n", "Line 0: Error read 3 floating points of float type"),
    ("e.obj", "v 0 0 # 1\n", "Line 0: Error read 3 floating points of float type"),
    ("e.obj", "vt 0.5 0.5 1\n", "Line 0: 3-dimensional textures are not supported"),
    ("e.obj", "vt 0.5\n", "Line 0: Error read 2 or 3 floating points of float type"),
    ("e.obj", "f 1/x 2 3\n", "Line 0: Error read facet vertex number"),
    ("e.obj", "f a 2 3\n", "Line 0: Error read facet vertex first number"),
    ("e.obj", "v 0 0 0\n", "No facets found in OBJ file"),
    ("e.obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n", "Vertex index 3 is out of bounds [0, 3)"),
    ("e.obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf -4 2 3\n", "Vertex index -1 is out of bounds [0, 3)"),
    ("e.obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1/1 2/1 3/1\n", "Texture coordinate index 0 is out of bounds [0, 0)"),
    ("e.obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1//2 2//1 3//1\n", "Normal index 1 is out of bounds [0, 1)"),
    ("e.obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n", "Object size error"),  # flat in z
    ("e.obj", "v 0 0 0\nv 1 1 1\nv 2 2 2\nf 1 2 3\n", "No 2-facets found in OBJ-3 file"),
    ("e.objx", "f 1 2 3\n", 'Wrong dimension number string "x"'),
    ("e.obj2", "f 1 2 3\n", "Wrong dimension number 2"),
    ("e.txt", "f 1 2 3\n", "Unsupported OBJ file format txt"),
    ("e", "f 1 2 3\n", "No OBJ file extension found"),
]


@pytest.mark.parametrize("name,text,msg", ERRORS)
def test_errors_match(tmp_path, name, text, msg):
    p = write_text(tmp_path, name, text)
    with pytest.raises(obj_ref.ObjError) as e1:
        obj_ref.read(p)
    assert str(e1.value) == msg
    with pytest.raises(native.NspError) as e2:
        native.read_obj(p, 2)
    assert msg in str(e2.value)


# ---------------------------------------------------------------------------
# Random files: parallel reader against the oracle

def random_obj(rng, N, lines, crlf=False, normals=True, tex=True):
    out = []
    nv = nvn = nvt = 0
    fmt = ["%.6g", "%.9e", "%+.4f", "%.1f"]
    max_groups = (5 if N == 3 else 1) + N - 1
    for _ in range(lines):
        r = rng.random()
        if r < 0.45 or nv < N:
            x = rng.normal(size=N + (1 if rng.random() < 0.1 else 0)) * 10.0 ** int(rng.integers(-3, 4))
            sep = " " if rng.random() < 0.8 else "\t "
            out.append("v" + "".join(sep + (fmt[rng.integers(len(fmt))] % c) for c in x))
            nv += 1
        elif r < 0.55 and normals:
            x = rng.normal(size=N) + 0.01
            out.append("vn " + " ".join("%.7g" % c for c in x))
            nvn += 1
        elif r < 0.6 and tex:
            x = rng.random(size=N - 1)
            out.append("vt " + " ".join("%.5f" % c for c in x) + (" 0" if rng.random() < 0.3 else ""))
            nvt += 1
        elif r < 0.67:
            out.append(["# comment", "o part", "g group", "usemtl m", "s 1", "", "   "][rng.integers(7)])
        else:
            G = int(rng.integers(N, max_groups + 1))
            kind = int(rng.integers(4))
            if kind in (1, 3) and nvt == 0:
                kind -= 1
            if kind in (2, 3) and nvn == 0:
                kind = 0
            groups = []
            for _ in range(G):
                vi = int(rng.integers(nv))
                s = str(vi + 1) if rng.random() < 0.6 else str(vi - nv)
                ti, ni = int(rng.integers(max(nvt, 1))), int(rng.integers(max(nvn, 1)))
                t = str(ti + 1) if rng.random() < 0.5 else str(ti - nvt)
                n = str(ni + 1) if rng.random() < 0.5 else str(ni - nvn)
                s += ["", "/" + t, "//" + n, "/" + t + "/" + n][kind]
                groups.append(s)
            out.append("f " + " ".join(groups) + (" # tail" if rng.random() < 0.1 else ""))
    eol = "\r\n" if crlf else "\n"
    return eol.join(out) + eol


@pytest.mark.parametrize("N,seed", [(3, 0), (3, 1), (4, 2), (5, 3), (6, 4), (3, 5)])
def test_random_files_match_oracle(tmp_path, N, seed):
    rng = np.random.default_rng(seed)
    text = random_obj(rng, N, 600, crlf=(seed == 5))
    p = write_text(tmp_path, "r.obj" if N == 3 else f"r.obj{N}", text)
    ref = obj_ref.read(p)
    assert ref["facets"].shape[0] > 50
    for threads in (1, 3, 8):
        assert_same(lib_read(p, threads), ref)


def test_large_file_automatic_chunks(tmp_path):
    rng = np.random.default_rng(11)
    text = random_obj(rng, 4, 60000)
    assert len(text) > 2 << 20
    p = write_text(tmp_path, "big.obj4", text)
    ref = obj_ref.read(p)
    assert_same(lib_read(p, 0), ref)
    assert_same(lib_read(p, 5), ref)


@pytest.mark.parametrize("seed", range(6))
def test_random_corruptions_give_same_error(tmp_path, seed):
    # one line of a valid file replaced by a malformed variant: both readers fail
    # with the same message, and the line number survives the chunked read
    rng = np.random.default_rng(100 + seed)
    lines = random_obj(rng, 3, 400).split("\n")
    k = int(rng.integers(len(lines) // 2, len(lines) - 1))
    lines[k] = ["f 1 0 2", "v 1 2", "f 1//1 2 3", "vt 1 1 1", "f 1", "f 1/2/x 2 3"][seed]
    p = write_text(tmp_path, "c.obj", "\n".join(lines))
    with pytest.raises(obj_ref.ObjError) as e1:
        obj_ref.read(p)
    assert str(e1.value).startswith(f"Line {k}: ")
    for threads in (1, 4):
        with pytest.raises(native.NspError) as e2:
            native.read_obj(p, threads)
        assert str(e1.value) in str(e2.value)


def float_tokens(rng, count):
    """Decimal spellings of many kinds: random magnitudes and digit counts,
    signs, leading zeros, exponents, float midpoints (shortest and exact
    spellings), more than 19 digits, integers above 2^24."""
    from decimal import Decimal
    out = []
    for _ in range(count):
        k = int(rng.integers(9))
        if k == 0:
            x = float(rng.normal()) * 10.0 ** int(rng.integers(-30, 31))
            out.append(repr(x))
        elif k == 1:
            out.append("%.*e" % (int(rng.integers(0, 12)), float(rng.normal()) * 10.0 ** int(rng.integers(-20, 21))))
        elif k == 2:
            out.append("%.*f" % (int(rng.integers(0, 10)), float(rng.normal()) * 10.0 ** int(rng.integers(-3, 6))))
        elif k in (3, 4):
            f = np.float32(rng.normal() * 10.0 ** int(rng.integers(-10, 11)))
            g = np.nextafter(f, np.float32(np.inf))
            mid = (float(f) + float(g)) / 2
            out.append(repr(mid) if k == 3 else str(Decimal(mid)))
        elif k == 5:
            out.append(("+" if rng.random() < 0.5 else "-") + "0" * int(rng.integers(0, 4))
                       + str(int(rng.integers(1, 10 ** 9))) + "." + str(int(rng.integers(0, 10 ** 6))).zfill(6))
        elif k == 6:
            out.append("".join(str(int(d)) for d in rng.integers(0, 10, size=int(rng.integers(18, 30))))
                       + "e" + str(int(rng.integers(-40, 10))))
        elif k == 7:
            out.append(str(int(rng.integers(1 << 24, 1 << 40))))
        else:
            out.append(["0", "-0", "0.0", ".5", "5.", "-.25e1", "1E3", "7e+0", "000", "0.000000000"][int(rng.integers(10))])
    return out


@pytest.mark.parametrize("seed", range(3))
def test_float_spellings_read_as_strtof_reads_them(tmp_path, seed):
    # the library's float reading (an exact fast path with strtof deciding the
    # hard cases) against libc strtof itself, bit for bit
    rng = np.random.default_rng(seed)
    tok = float_tokens(rng, 30000)
    lines = ["v " + " ".join(tok[i:i + 3]) for i in range(0, len(tok) - 2, 3)]
    lines.append("f 1 2 3")
    lines.append("f 4 5 6")
    p = write_text(tmp_path, "fl.obj", "\n".join(lines) + "\n")
    ref = obj_ref.read(p)  # every spelling here is a finite normal float
    got = lib_read(p, 4)
    assert got["vertices"].tobytes() == ref["vertices"].tobytes()


@pytest.mark.parametrize("tok", ["1e", "1e+", "0x1p3", "inf", "nan", "1e39", "1e-39", "1e-46", "-1e-45", "+",
                                 "-", ".", "1.5abc", "--1", "3.4028235e38", "3.4028236e38", "1.17549435e-38",
                                 "1.1754942e-38", "16777217", "0.1", "1" + "0" * 30])
def test_float_edge_spellings(tmp_path, tok):
    # one token at a time, as the first coordinate: both readers agree on the
    # value or on the error
    text = f"v {tok} 1 1\nv 0 0 2\nv 0 2 0\nv 3 0 0\nf 2 3 4\n"
    p = write_text(tmp_path, "edge.obj", text)
    try:
        ref = obj_ref.read(p)
    except obj_ref.ObjError as e:
        with pytest.raises(native.NspError) as e2:
            native.read_obj(p, 1)
        assert str(e) in str(e2.value)
        return
    assert lib_read(p, 1)["vertices"].tobytes() == ref["vertices"].tobytes()


# ---------------------------------------------------------------------------
# Writer

@pytest.mark.parametrize("N,with_normals", [(3, True), (3, False), (4, False), (5, True)])
def test_writer_matches_oracle_and_round_trips(tmp_path, N, with_normals):
    rng = np.random.default_rng(N)
    v, f = native.meshgen_cube_sphere(N, 2 if N > 3 else 4, True, 0.03, 0.0, 7)
    v = v * np.float32(3.7) + np.float32(1.25)
    nrm = fn = None
    if with_normals:
        nrm = (v + rng.normal(size=v.shape).astype(np.float32) * 0.1).astype(np.float32)
        fn = f.copy()
        fn[::5] = -1  # some facets without normals
    ext = "obj" if N == 3 else f"obj{N}"
    p1 = os.path.join(str(tmp_path), "a." + ext)
    p2 = os.path.join(str(tmp_path), "b." + ext)
    text = obj_ref.write(p1, v, f, nrm, fn, comment="round trip")
    native.write_obj(p2, v, f, nrm, fn, comment="round trip")
    assert open(p2).read() == text
    # the reference's own check (src/geometry/test/test_reconstruction.cpp:88-123): counts survive
    back = lib_read(p2)
    assert_same(back, obj_ref.read(p2))
    assert back["vertices"].shape == v.shape and back["facets"].shape[0] == f.shape[0]
    assert back["normals"].shape[0] == (0 if nrm is None else nrm.shape[0])
    assert np.abs(back["vertices"]).max() <= 1.0 + 1e-6


def test_writer_errors(tmp_path):
    v = np.zeros((3, 3), np.float32)
    f = np.array([[0, 1, 2]], np.int32)
    with pytest.raises(native.NspError, match="equal to each other"):
        native.write_obj(os.path.join(str(tmp_path), "z.obj"), v, f)
    v[1, 0] = 1
    with pytest.raises(native.NspError, match="extension"):
        native.write_obj(os.path.join(str(tmp_path), "z.obj4"), v, f)
    with pytest.raises(native.NspError, match="out of bounds"):
        native.write_obj(os.path.join(str(tmp_path), "z.obj"), v, f + 1)


# ---------------------------------------------------------------------------
# GPU: a mesh from a file equals the mesh from the same arrays

@pytest.mark.gpu
def test_mesh_from_obj_equals_mesh_from_arrays(ctx, tmp_path):
    v, f = native.meshgen_cube_sphere(4, 3, True, 0.02, 0.0, 3)
    p = os.path.join(str(tmp_path), "m.obj4")
    native.write_obj(p, v, f)
    obj = native.read_obj(p)
    print(f"obj mesh V={len(obj.vertices)} F={len(obj.facets)}", flush=True)
    a = native.Mesh.from_obj(ctx, obj)
    b = ctx.mesh(obj.vertices, obj.facets)
    for x, y in zip(a.records(), b.records()):
        assert np.array_equal(x, y)
    assert np.array_equal(a.tree_dump(), b.tree_dump())
    from math_amd import configs
    cfg = configs.scene_dict(obj.vertices, obj.facets, 10, (2.0,) * 4)
    imgs = []
    for m in (a, b):
        sc = native.Scene(ctx, m, cfg)
        sc.paint(4, 3, seed=9)
        imgs.append(sc.image())
    assert np.array_equal(imgs[0][0], imgs[1][0])
