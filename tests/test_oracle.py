"""CPU tests of the FP64 checker: known answers, closed forms, invariants.

The reference holds no golden vectors for the painter path (its tests seed from
/dev/urandom), so the checker is pinned by its known-answer tests (cosine
sphere coefficient static asserts), published vectors (Philox4x32-10), closed
forms (barycentric identities, orthonormal complements, the sampler count
rule), brute force (tree trace == linear scan) and the reference's own
closed-mesh invariant (src/painter/shapes/test/test_mesh.cpp:106-223).
"""
import math

import numpy as np
import pytest

from math_amd import native
from oracle import oracle as O


def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32_10
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_cosine_sphere_coefficient_known_answers():
    # static_asserts of src/painter/coefficient/cosine_sphere.cpp:35-51
    kat = {2: math.pi / 2, 3: 2.0, 4: 3 * math.pi / 4, 5: 8 / 3, 6: 15 * math.pi / 16, 7: 16 / 5,
           8: 35 * math.pi / 32, 9: 128 / 35, 10: 315 * math.pi / 256, 15: 2048 / 429,
           20: 230945 * math.pi / 131072, 25: 4194304 / 676039}
    for n, v in kat.items():
        assert O.cosine_sphere_coefficient(n) == pytest.approx(v, rel=1e-15)
    # closed form B(1/2,(n-1)/2)/B(1,(n-1)/2), test_cosine_sphere.cpp:47-84
    for n in range(2, 40):
        a, b = 0.5, (n - 1) / 2
        beta1 = math.exp(math.lgamma(a) + math.lgamma(b) - math.lgamma(a + b))
        beta2 = math.exp(math.lgamma(1) + math.lgamma(b) - math.lgamma(1 + b))
        assert O.cosine_sphere_coefficient(n) == pytest.approx(beta1 / beta2, rel=1e-12)


@pytest.mark.parametrize("dims,spp,expect", [(2, 16, 4), (3, 16, 3), (3, 4, 2), (3, 1024, 11), (3, 4096, 16),
                                             (4, 8, 2), (4, 4, 2), (2, 1, 1), (3, 1, 1), (2, 17, 5)])
def test_sampler_count_rule(dims, spp, expect):
    # StratifiedJitteredSampler::one_dimension_size (sampler.h:31-52): smallest k with k^dims >= spp
    assert O.samples_per_dimension(dims, spp) == expect


@pytest.mark.parametrize("n", [3, 4, 5, 6])
def test_orthogonal_complement_is_orthonormal(n):
    rng = np.random.default_rng(n)
    for _ in range(200):
        u = rng.normal(size=n)
        u /= np.linalg.norm(u)
        if rng.random() < 0.3:
            u = np.zeros(n)
            u[rng.integers(n)] = rng.choice([-1.0, 1.0])
        B = O.orthogonal_complement(u)
        M = np.vstack([u, B])
        assert np.allclose(M @ M.T, np.eye(n), atol=1e-12)


@pytest.mark.parametrize("n", [3, 4, 5, 6])
def test_ortho_nn_is_cofactor_vector(n):
    rng = np.random.default_rng(10 + n)
    for _ in range(50):
        V = rng.normal(size=(n - 1, n))
        w = O.ortho_nn(V)
        assert np.allclose(V @ w, 0, atol=1e-10)
        # |w| = (n-1)-volume of the parallelotope = sqrt(det(V V^T))
        assert np.linalg.norm(w) == pytest.approx(math.sqrt(np.linalg.det(V @ V.T)), rel=1e-9)
        # orientation: det([V; w]) > 0 when w = sum_i (-1)^i minor_i e_i with the sign of the last row
        assert np.linalg.det(np.vstack([V, w])) * (-1) ** (n - 1) > 0


@pytest.mark.parametrize("n", [3, 4, 5, 6])
def test_cosine_hemisphere_distribution(n):
    # test_sphere.cpp:69-123: histogram of cos(theta) is proportional to cos
    rng = np.random.default_rng(20 + n)
    normal = rng.normal(size=n)
    normal /= np.linalg.norm(normal)
    cs = []
    for _ in range(20000):
        d = O.cosine_hemisphere(normal, rng.random(8))
        assert abs(np.linalg.norm(d) - 1) < 1e-12
        cs.append(d @ normal)
    cs = np.array(cs)
    assert cs.min() > 0
    hist, edges = np.histogram(cs, bins=10, range=(0, 1))
    centers = (edges[:-1] + edges[1:]) / 2
    expect = 2 * centers * 0.1 * len(cs)
    assert np.all(np.abs(hist - expect) / expect < 0.12)


def _mesh(n, g, sigma=0.02, cap=0.0, seed=1):
    return native.meshgen_cube_sphere(n, g, True, sigma, cap, seed)


@pytest.mark.parametrize("n,g", [(3, 6), (4, 3), (5, 2)])
def test_barycentric_records_closed_form(n, g):
    v, f = _mesh(n, g)
    m = O.Mesh(v, f)
    nrm, pn, pd, v0 = m.records()
    V = v.astype(np.float64)[f]  # F, n(vertices), n
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1, atol=1e-14)
    # normal orthogonal to every edge
    edges = V[:, 1:, :] - V[:, :1, :]
    assert np.max(np.abs(np.einsum("fkn,fn->fk", edges, nrm))) < 1e-12
    # barycentric coordinate i of vertex j is delta_ij (i < n-1); plane normals orthogonal to the simplex normal
    bc = np.einsum("fvn,fin->fvi", V, pn) - pd[:, None, :]
    eye = np.zeros((n, n - 1))
    eye[: n - 1] = np.eye(n - 1)
    assert np.max(np.abs(bc - eye[None])) < 1e-9
    assert np.max(np.abs(np.einsum("fin,fn->fi", pn, nrm))) < 1e-9


@pytest.mark.parametrize("n,g", [(3, 8), (4, 4), (5, 2)])
def test_tree_trace_equals_linear_scan(n, g):
    v, f = _mesh(n, g)
    m = O.Mesh(v, f)
    m.build_tree()
    rng = np.random.default_rng(n)
    K = 1500
    org = rng.normal(size=(K, n))
    org = org / np.linalg.norm(org, axis=1, keepdims=True) * 3
    d = rng.normal(size=(K, n)) * 0.5 - org
    i1, t1, near = m.trace_brute(org.astype(np.float32), d.astype(np.float32))
    i2, t2 = m.trace_tree(org.astype(np.float32), d.astype(np.float32))
    assert np.array_equal(i1, i2)
    assert np.array_equal(t1, t2)
    assert (i1 >= 0).sum() > K // 3


@pytest.mark.parametrize("n,g", [(3, 10), (4, 4)])
def test_closed_mesh_enter_exit_miss(n, g):
    # test_mesh.cpp:106-223: a ray into a closed convex mesh hits, hits again after
    # moving past the first hit, then misses; at most 0.05% failures
    v, f = _mesh(n, g, sigma=0.0)
    m = O.Mesh(v, f)
    m.build_tree()
    rng = np.random.default_rng(3)
    K = 2000
    org = rng.normal(size=(K, n))
    org = org / np.linalg.norm(org, axis=1, keepdims=True) * 4
    tgt = rng.normal(size=(K, n))
    tgt *= (0.6 * rng.random(K) / np.linalg.norm(tgt, axis=1))[:, None]  # aim inside the inscribed ball
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    i1, t1 = m.trace_tree(org.astype(np.float32), d.astype(np.float32))
    hit = i1 >= 0
    o2 = org + d * (t1[:, None] + 1e-6)
    i2, t2 = m.trace_tree(o2.astype(np.float32), d.astype(np.float32))
    o3 = o2 + d * (t2[:, None] + 1e-6)
    i3, _ = m.trace_tree(o3.astype(np.float32), d.astype(np.float32))
    errors = np.sum(~hit) + np.sum(hit & (i2 < 0)) + np.sum(hit & (i2 >= 0) & (i3 >= 0))
    assert errors <= max(1, 0.0005 * K)


def test_tree_leaf_rule_and_containment():
    # every leaf holds <= 10 objects unless at the maximum depth (tree.h:355-358), and every
    # simplex whose sampled interior points lie in a leaf box is listed in that leaf
    n, g = 4, 3
    v, f = _mesh(n, g)
    m = O.Mesh(v, f)
    m.build_tree()
    dump = m.tree_dump()
    mn, mx = m.min_max()
    eps = 2.220446049250313e-16
    absmax = np.maximum(np.abs(mn), np.abs(mx))
    lo = mn - absmax * 20 * eps
    size = (mx + absmax * 20 * eps) - lo
    leaves = {}
    k = 0
    while k < len(dump):
        lv, q, c = dump[k], tuple(dump[k + 1:k + 1 + n]), dump[k + 1 + n]
        ids = dump[k + 2 + n:k + 2 + n + c]
        leaves[(lv, q)] = set(ids.tolist())
        assert c <= 10 or lv == 8 - 1
        k += 2 + n + c
    rng = np.random.default_rng(5)
    V = v.astype(np.float64)[f]
    for _ in range(3000):
        s = rng.integers(len(f))
        w = rng.dirichlet(np.ones(n))
        p = w @ V[s]
        # find the leaf containing p by descending levels
        found = False
        for lv in range(0, 8):
            cell = size / 2 ** lv
            q = tuple(np.floor((p - lo) / cell).astype(int).tolist())
            if (lv, q) in leaves:
                assert s in leaves[(lv, q)]
                found = True
                break
        assert found


# ---------------------------------------------------------------------------
# Projectors (src/painter/projectors/projector.h): closed forms of the ray directions

def _proj_cfg(n, kind, angle, screen, upp=0.01):
    cfg = dict(n=n, camera_org=[0.3] * n, camera_dir=[0.0] * (n - 1) + [-2.0],
               axes=[[3.0 if j == i else 0.0 for j in range(n)] for i in range(n - 1)],
               screen=list(screen), projector=kind, view_angle=angle, units_per_pixel=upp,
               light_pos=[1.0] * n, light_rgb=[1, 1, 1], light_unit_distance=1.0, background_rgb=[0, 0, 0],
               object_rgb=[1, 1, 1], diffuse=0.8)
    return cfg


@pytest.mark.parametrize("n", [3, 4, 5])
def test_projector_ray_closed_forms(n):
    rng = np.random.default_rng(n)
    screen = [int(x) for x in rng.integers(20, 60, size=n - 1)]
    w = screen[0]
    for _ in range(20):
        point = rng.random(n - 1) * screen
        sp = point - np.array(screen) / 2
        # parallel: origin moves on the screen plane, direction fixed
        org, d = O.projector_ray(_proj_cfg(n, 0, 0.0, screen), point)
        assert np.allclose(org[:n - 1], 0.3 + 0.01 * sp, atol=1e-12) and org[n - 1] == 0.3
        assert np.allclose(d, [0] * (n - 1) + [-1])
        # perspective: the screen at distance L = (w/2) / tan(angle/2) in pixel units
        ang = float(rng.uniform(20, 150))
        org, d = O.projector_ray(_proj_cfg(n, 1, ang, screen), point)
        L = (w / 2) / np.tan(np.radians(ang) / 2)
        expect = np.concatenate([sp, [-L]]) / np.sqrt(sp @ sp + L * L)
        assert np.allclose(org, 0.3) and np.allclose(d, expect, atol=1e-12)
        # spherical: the screen point lifted onto the sphere of radius R = (w/2) / sin(angle/2)
        ang = float(rng.uniform(10, 60))
        R = (w / 2) / np.sin(np.radians(ang) / 2)
        if sp @ sp < R * R and (np.sin(np.radians(ang) / 2) ** 2) * (1 + sum((s / w) ** 2 for s in screen[1:])) < 1:
            org, d = O.projector_ray(_proj_cfg(n, 2, ang, screen), point)
            expect = np.concatenate([sp, [-np.sqrt(R * R - sp @ sp)]]) / R
            assert np.allclose(d, expect, atol=1e-12)
    # the edge of the screen is seen at half the view angle
    for kind in (1, 2):
        _, d = O.projector_ray(_proj_cfg(n, kind, 50.0, screen), [0.0] + [s / 2 for s in screen[1:]])
        assert np.isclose(np.degrees(np.arccos(-d[n - 1])), 25.0)


def test_projector_errors():
    with pytest.raises(ValueError):
        O.projector_ray(_proj_cfg(3, 1, 180.0, [10, 10]), [1, 1])
    with pytest.raises(ValueError):
        O.projector_ray(_proj_cfg(3, 2, 170.0, [10, 10]), [1, 1])  # too big for a square screen
    with pytest.raises(ValueError):
        O.projector_ray(_proj_cfg(3, 0, 0.0, [10, 10], upp=0.0), [1, 1])


# ---------------------------------------------------------------------------
# Shapes (src/painter/space/parallelotope.h, hyperplane_parallelotope.h): the
# oracle's plane-based tests against the parameter form p = org + E c, solved
# by linear algebra

@pytest.mark.parametrize("n", [3, 4, 5])
def test_parallelotope_intersection_by_coordinates(n):
    rng = np.random.default_rng(40 + n)
    checked = 0
    for _ in range(300):
        E = rng.normal(size=(n, n)) + 2 * np.eye(n)
        org = rng.normal(size=n)
        o = org + rng.normal(size=n) * 3
        d = (org + E.sum(axis=0) * rng.random() - o)
        d /= np.linalg.norm(d)
        Einv = np.linalg.inv(E.T)          # c = Einv (p - org), columns of E.T are the edges
        c0, c1 = Einv @ (o - org), Einv @ d
        lo, hi = -np.inf, np.inf
        for i in range(n):
            if c1[i] == 0:
                continue
            a, b = (0 - c0[i]) / c1[i], (1 - c0[i]) / c1[i]
            lo, hi = max(lo, min(a, b)), min(hi, max(a, b))
        hit = hi > max(lo, 0)
        r = O.shape_intersect(0, org, E, o, d)
        if not hit or min(abs(hi - lo), abs(lo), abs(hi)) < 1e-9:
            assert (r is None) == (not hit) or min(abs(hi - lo), abs(lo), abs(hi)) < 1e-9
            continue
        assert r is not None
        t, nrm = r
        expect = lo if lo > 0 else hi
        assert abs(t - expect) < 1e-9 * (1 + abs(expect))
        # the normal of the face the hit lies on: the gradient of the coordinate that is 0 or 1 there
        c = c0 + t * c1
        i = int(np.argmin(np.minimum(np.abs(c), np.abs(c - 1))))
        g = Einv[i] * (1 if abs(c[i] - 1) < abs(c[i]) else -1)
        assert np.allclose(nrm, g / np.linalg.norm(g), atol=1e-9)
        checked += 1
    assert checked > 100


@pytest.mark.parametrize("n", [3, 4, 5])
def test_hyperplane_parallelotope_intersection_by_coordinates(n):
    rng = np.random.default_rng(60 + n)
    hits = misses = 0
    for _ in range(300):
        E = rng.normal(size=(n - 1, n))
        org = rng.normal(size=n)
        o = org + rng.normal(size=n) * 3
        target = org + E.T @ rng.uniform(-0.3, 1.3, size=n - 1)
        d = (target - o) * (1 if rng.random() < 0.8 else -1)  # some rays point away
        d /= np.linalg.norm(d)
        # org + E^T c = o + t d  ->  [E^T, -d] [c; t] = o - org
        sol = np.linalg.solve(np.column_stack([E.T, -d]), o - org)
        c, t = sol[:-1], sol[-1]
        inside = t > 0 and np.all(c > 0) and np.all(c < 1)
        if np.min(np.abs(np.concatenate([c, c - 1, [t]]))) < 1e-9:
            continue
        r = O.shape_intersect(1, org, E, o, d)
        assert (r is not None) == inside
        if inside:
            assert abs(r[0] - t) < 1e-9 * (1 + abs(t))
            assert np.allclose(E @ r[1], 0, atol=1e-9) and abs(np.linalg.norm(r[1]) - 1) < 1e-12
            hits += 1
        else:
            misses += 1
    assert hits > 20 and misses > 20


# ---------------------------------------------------------------------------
# Several light sources (direct_diffuse_lighting's loop, painter.cpp:205-289)

@pytest.mark.parametrize("n", [3, 4])
def test_lights_add_linearly(n):
    # without bounces a hit's colour is the sum of the lights' terms, so the render
    # with two lights equals the two single-light renders minus the unlit one
    from math_amd import configs
    g = {3: 6, 4: 3}[n]
    v, f = native.meshgen_cube_sphere(n, g, True, 0.02, 0.0, 5)
    cfg = configs.scene_dict(v, f, 8 if n == 4 else 16, tuple([2.0] * n))
    la = dict(kind=0, pos=[2.0] * n, rgb=[1.0, 0.8, 0.6], unit_distance=2.0)
    lb = dict(kind=1, pos=[-1.5] + [1.0] * (n - 1), rgb=[0.2, 0.3, 0.9])
    m = O.Mesh(v, f)
    m.build_tree()

    def render(lights):
        c = dict(cfg)
        c["lights"] = lights
        img, rays = O.render(m, O.make_scene(c, 4, 0, 7), threads=4)
        return img, rays
    ab, rab = render([la, lb])
    a, ra = render([la])
    b, rb = render([lb])
    z, rz = render([])
    assert np.allclose(ab, a + b - z, rtol=1e-5, atol=1e-6)
    # each light's shadow ray counts once: the ray counts add the same way
    assert rab == ra + rb - rz
    assert not np.allclose(a, z) and not np.allclose(b, z)


@pytest.mark.parametrize("n,g", [(3, 12), (4, 4), (5, 2)])
def test_tree_build_threads_equal_one_thread(n, g):
    # the subtrees built by worker threads give the one-thread tree: same boxes,
    # leaf lists and traced hits (only the node numbering differs)
    v, f = _mesh(n, g)
    a, b = O.Mesh(v, f), O.Mesh(v, f)
    a.build_tree(threads=1)
    b.build_tree(threads=6)
    assert a.tree_info() == b.tree_info()

    def leaves(m):
        d, out, k = m.tree_dump(), {}, 0
        while k < len(d):
            c = int(d[k + 1 + n])
            out[tuple(int(x) for x in d[k:k + 1 + n])] = d[k + 2 + n:k + 2 + n + c].tolist()
            k += 2 + n + c
        return out
    assert leaves(a) == leaves(b)
    rng = np.random.default_rng(n)
    org = rng.normal(size=(500, n))
    org = (org / np.linalg.norm(org, axis=1, keepdims=True) * 3).astype(np.float32)
    d = (rng.normal(size=(500, n)) * 0.5 - org).astype(np.float32)
    ia, ta = a.trace_tree(org, d)
    ib, tb = b.trace_tree(org, d)
    assert np.array_equal(ia, ib) and np.array_equal(ta, tb)
