"""Convex hull and Delaunay simplices (compute_convex_hull / compute_delaunay,
geometry/core/convex_hull.cpp).

CPU tests pin the oracle (oracle/hull_ref.py):
- against Qhull (scipy.spatial) facet sets on points in general position;
- against the cross-polytope's closed form;
- against the checks of the reference's own test (test_convex_hull.cpp:112-141);
- against the empty-circumsphere property of Delaunay simplices.
GPU tests compare the library (exact predicates, GPU conflict pass) with the
oracle run in the library's own insertion order.  The vertex tuples must be
equal, degenerate inputs included, and the unit normals equal to within 1e-12.
"""
import os
import subprocess

import numpy as np
import pytest
from scipy.spatial import ConvexHull, Delaunay

from math_amd import configs, native
from oracle import hull_ref as H

HERE = os.path.dirname(os.path.abspath(__file__))


def ball_points(n, count, seed, on_sphere=False, zero=False):
    return configs.hull_points(n, count, seed, on_sphere, zero)


def ridges_two_facets(vertices):
    cnt = {}
    for f in vertices:
        for r in range(len(f)):
            key = tuple(sorted(f[:r] + f[r + 1:]))
            cnt[key] = cnt.get(key, 0) + 1
    return all(c == 2 for c in cnt.values())


def nothing_beyond(points, facets, eps=1e-5):
    # test_convex_hull.cpp:97-108 allows 0.01 (CHECK_EPSILON); the 2^-30 grid of
    # the integer points leaves about 1e-6 here
    p = np.asarray(points, np.float64)
    for v, o in facets:
        d = p - p[v[0]]
        nrm = np.linalg.norm(d, axis=1)
        nrm[nrm == 0] = 1
        if np.max((d @ np.asarray(o)) / nrm) > eps:
            return False
    return True


# ---------------------------------------------------------------- oracle (CPU)

@pytest.mark.parametrize("n,count", [(2, 200), (3, 200), (4, 120), (5, 50)])
def test_oracle_hull_matches_qhull(n, count):
    p = ball_points(n, count, 10 + n)
    order = np.random.default_rng(n).permutation(count)
    ref = H.convex_hull(p, order)
    q = ConvexHull(p.astype(np.float64))
    assert set(frozenset(v) for v, _ in ref) == set(frozenset(s) for s in q.simplices.tolist())
    # Qhull's outward unit normals
    qn = {frozenset(s): e[:n] for s, e in zip(q.simplices.tolist(), q.equations)}
    for v, o in ref:
        assert np.allclose(o, qn[frozenset(v)], atol=1e-6)


@pytest.mark.parametrize("n", [2, 3, 4, 5])
def test_oracle_cross_polytope(n):
    # +-e_i and interior points: 2^n facets, one per sign pattern, normal s / sqrt(n)
    pts = []
    for i in range(n):
        for s in (1.0, -1.0):
            e = np.zeros(n)
            e[i] = s
            pts.append(e)
    rng = np.random.default_rng(n)
    inner = rng.uniform(-1, 1, (40, n))
    inner = inner[np.abs(inner).sum(axis=1) < 0.9]
    p = np.vstack([np.array(pts), inner]).astype(np.float32)
    order = rng.permutation(len(p))
    ref = H.convex_hull(p, order)
    assert len(ref) == 2 ** n
    for v, o in ref:
        assert all(k < 2 * n for k in v)
        signs = np.zeros(n)
        for k in v:
            signs[k // 2] = 1.0 if k % 2 == 0 else -1.0
        assert np.all(signs != 0)
        assert np.allclose(o, signs / np.sqrt(n), atol=1e-12)


@pytest.mark.parametrize("n,zero", [(3, False), (3, True), (4, False), (4, True)])
def test_oracle_reference_checks(n, zero):
    # check_convex_hull (test_convex_hull.cpp:112-141)
    p = ball_points(n, 150, 20 + n, zero=zero)
    ref = H.convex_hull(p, np.arange(len(p)))
    assert ridges_two_facets([list(v) for v, _ in ref])
    assert nothing_beyond(p, ref)


@pytest.mark.parametrize("n,count", [(2, 120), (3, 60)])
def test_oracle_delaunay_matches_qhull_and_empty_spheres(n, count):
    p = np.random.default_rng(30 + n).random((count, n)).astype(np.float32)
    order = np.random.default_rng(1).permutation(count)
    ref, ints, kept = H.delaunay(p, order)
    q = Delaunay(p.astype(np.float64))
    assert set(frozenset(v) for v, _ in ref) == set(frozenset(s) for s in q.simplices.tolist())
    # empty circumsphere on the integer points (exact in float64 at these sizes: relative check)
    X = np.zeros((count, n))
    for i, k in enumerate(kept):
        X[k] = ints[i]
    for v, normals in ref:
        a = X[list(v)]
        A = 2 * (a[1:] - a[0])
        b = (a[1:] ** 2).sum(1) - (a[0] ** 2).sum()
        c = np.linalg.solve(A, b)
        r2 = ((a[0] - c) ** 2).sum()
        d2 = ((X - c) ** 2).sum(1)
        assert np.all(d2 >= r2 * (1 - 1e-9))
        # each normal points away from its vertex
        for r, o in enumerate(normals):
            face = [v[i] for i in range(n + 1) if i != r]
            assert np.dot(np.asarray(o), X[v[r]] - X[face[0]]) < 0


def test_oracle_errors():
    with pytest.raises(ValueError):
        H.convex_hull(np.ones((10, 3), np.float32), np.arange(10))
    flat = np.random.default_rng(0).random((20, 3)).astype(np.float32)
    flat[:, 2] = 0.5
    with pytest.raises(ValueError):
        H.convex_hull(flat, np.arange(20))


# ---------------------------------------------------------------- library (GPU)

def same_as_oracle(h, p, delaunay=False):
    if delaunay:
        ref, _, _ = H.delaunay(p, h.order)
    else:
        ref = H.convex_hull(p, h.order)
    got = {tuple(v): o for v, o in zip(h.vertices.tolist(), h.orthos)}
    exp = {v: np.asarray(o) for v, o in ref}
    assert set(got) == set(exp)
    for k in exp:
        assert np.max(np.abs(np.asarray(got[k]) - exp[k])) < 1e-12
    return len(exp)


@pytest.mark.gpu
@pytest.mark.parametrize("n,count,first", [(2, 300, 8), (3, 300, 16), (4, 200, 16), (4, 200, 0), (5, 70, 12),
                                           (6, 36, 10)])
def test_hull_matches_oracle(ctx, n, count, first):
    p = ball_points(n, count, 40 + n)
    print(f"hull n={n} count={count} first={first}", flush=True)
    h = native.convex_hull(ctx, p, first)
    k = same_as_oracle(h, p)
    assert len(h.order) == count
    print(f"facets {k} dropped {h.dropped}", flush=True)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [3, 4])
def test_hull_degenerate_matches_oracle(ctx, n):
    # the reference test's flat case: points in x_n = 0 and one apex; the base
    # is triangulated as the insertion order makes it
    p = ball_points(n, 160, 50 + n, zero=True)
    h = native.convex_hull(ctx, p, 12)
    same_as_oracle(h, p)
    assert ridges_two_facets(h.vertices.tolist())


@pytest.mark.gpu
@pytest.mark.parametrize("n,count", [(2, 200), (3, 100), (4, 50), (5, 28)])
def test_delaunay_matches_oracle(ctx, n, count):
    p = np.random.default_rng(60 + n).random((count, n)).astype(np.float32)
    h = native.delaunay(ctx, p, 10)
    same_as_oracle(h, p, delaunay=True)
    # the integer points the simplices were computed on
    ints, kept = H.quantise(p, h.order, (1 << 24) - 1)
    for i, k in enumerate(kept):
        assert np.array_equal(h.points[k], np.asarray(ints[i], np.float64))


@pytest.mark.gpu
def test_hull_gpu_pass_equals_host_only(ctx):
    # 20,000 points in the 4-ball: the GPU pass after 1024 points drops most of
    # them (about 14,000); the facets equal those of the host-only construction
    # and Qhull's
    p = ball_points(4, 20000, 70)
    a = native.convex_hull(ctx, p, 1024)
    b = native.convex_hull(ctx, p, 20000)
    print(f"dropped {a.dropped} of 20000, facets {len(a.vertices)}, ms {a.ms} vs host-only {b.ms}", flush=True)
    assert a.dropped > 12000 and b.dropped == 0
    assert set(map(tuple, a.vertices.tolist())) == set(map(tuple, b.vertices.tolist()))
    q = ConvexHull(p.astype(np.float64))
    assert set(frozenset(s) for s in a.vertices.tolist()) == set(frozenset(s) for s in q.simplices.tolist())
    assert ridges_two_facets(a.vertices.tolist())


@pytest.mark.gpu
def test_hull_on_sphere_properties(ctx):
    # every point of the sphere is a vertex
    p = ball_points(4, 3000, 71, on_sphere=True)
    h = native.convex_hull(ctx, p, 128)
    assert h.dropped == 0
    assert len(set(h.vertices.ravel().tolist())) == 3000
    assert ridges_two_facets(h.vertices.tolist())
    assert nothing_beyond(p, list(zip(h.vertices.tolist(), h.orthos)))


@pytest.mark.gpu
def test_hull_errors(ctx):
    with pytest.raises(native.NspError):
        native.convex_hull(ctx, np.ones((10, 3), np.float32))
    with pytest.raises(native.NspError):
        native.convex_hull(ctx, np.zeros((3, 3), np.float32) + np.eye(3, dtype=np.float32))
    with pytest.raises(native.NspError):
        native.convex_hull(ctx, np.random.default_rng(0).random((20, 7)).astype(np.float32))
    with pytest.raises(native.NspError):
        native.delaunay(ctx, np.random.default_rng(0).random((20, 6)).astype(np.float32))
    bad = np.random.default_rng(0).random((20, 3)).astype(np.float32)
    bad[4, 1] = np.nan
    with pytest.raises(native.NspError):
        native.convex_hull(ctx, bad)
    flat = np.random.default_rng(0).random((20, 3)).astype(np.float32)
    flat[:, 2] = 0.5
    with pytest.raises(native.NspError):
        native.convex_hull(ctx, flat)


@pytest.mark.gpu
def test_cpp_hull_order():
    exe = os.path.join(HERE, "cpp", "test_hull_order")
    assert os.path.exists(exe), "built by __graft_entry__.build()"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout


@pytest.mark.gpu
def test_sphere_hull_mesh_renders_as_checker(ctx):
    # the painter's convex-hull test mesh (sphere_mesh.cpp): hull of random
    # points on the 3-sphere, rendered in FP64 against the checker's render
    from oracle import oracle as O
    v, f = configs.sphere_hull_mesh(ctx, 4, 300, seed=3)
    assert len(f) > 300 and ridges_two_facets(f.tolist())
    cfg = configs.scene_dict(v, f, 8, (2.0, 2.0, 2.0, 2.0))
    gm = ctx.mesh(v, f)
    sc = native.Scene(ctx, gm, cfg)
    sc.paint(4, 3, precision=native.PRECISION_FP64, seed=4)
    img, cnt = sc.image()
    om = O.Mesh(v, f)
    om.build_tree()
    ref, rays = O.render(om, O.make_scene(cfg, 4, 3, 4), threads=8)
    rmse = float(np.sqrt(np.mean((img / cnt[:, None] - ref / cnt[:, None]) ** 2)))
    print(f"sphere hull mesh: {len(f)} facets, rmse {rmse:.2e}", flush=True)
    assert rmse < 1e-4
