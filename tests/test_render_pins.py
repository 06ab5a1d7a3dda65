"""Closed-form pins of the checker's composed render (orc_render).

Every GPU render test compares against orc_render, so its composition of
trace_path and direct_diffuse_lighting (src/painter/painter.cpp:197-397) is
pinned here by scenes whose colour has a closed form per sample, written from
the reference's formulas and not from the checker's code:

  * PointLight::properties (src/painter/lights/light_source.h:42-56): colour
    times u^(N-1) / r^(N-1); ConstantLight (:70-76): the colour itself;
  * the light weight DIFFUSE_LIGHT_COEFFICIENT<N> * dot(l, n) (painter.cpp:286-288)
    with the light skipped when dot <= 0 (:218-224) or its colour is below
    MIN_COLOR_LEVEL (:136-139, :211-214);
  * surface colour diffuse * rgb times (direct + diffuse bounce) (:384-393),
    cut off below MIN_COLOR_LEVEL (:384-387);
  * a bounce that leaves the scene returns the background's light colour, its
    luminance (painter.cpp:333-340, single_object.h:70-72), and a bounce that
    meets an emitting surface returns its light colour (:355-358);
  * the geometric normal turned towards the ray (:368-375);
  * the ray counts: one per trace_path call, one per object tested for a
    shadow ray until one blocks (:172-186).

The scenes are a hyperplane parallelotope facing a parallel camera (and, for
the bounce, an emitting plane above it), with the mesh object far behind the
plane, for n = 3..6.  The samples' screen points come from orc_sample_point,
whose stratification is pinned first (engine.h:86-156: sample k lies in the
cell of its base-k digits, dimension 0 least significant).  Also pinned here:
the single_object_scene conventions of math_amd.configs against
single_object.h:102-179, and the BarPaintbrush pixel order of the C++ drop-in
against paintbrush.h:33-110, 156-164.
"""
import itertools
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
COEF = {3: 2.0, 4: 3 * math.pi / 4, 5: 8 / 3, 6: 15 * math.pi / 16}  # cosine_sphere.cpp:35-51
SCREENS = {3: [7, 5], 4: [5, 4, 3], 5: [4, 3, 3, 2], 6: [3, 3, 2, 2, 2]}
CAM_Z = 5.0
UPP = 0.07


def f32(x):
    return np.float32(x)


def far_mesh(n, z=-10.0, shift=0.0):
    """One (n-1)-simplex in the hyperplane x_{n-1} = z, far behind the plane
    (and moved `shift` along x_0)."""
    v = np.zeros((n, n), np.float32)
    v[:, n - 1] = z
    v[:, 0] = shift
    for i in range(1, n):
        v[i, i - 1] += 1.0
    return v, np.arange(n, dtype=np.int32)[None]


def plane(n, z, half=50.0, rgb=(0.9, 0.6, 0.3), diffuse=0.8, light_rgb=None, flip=False):
    """Hyperplane parallelotope x_{n-1} = z spanning [-half, half]^(n-1)."""
    org = [-half] * (n - 1) + [z]
    edges = [[2 * half if j == i else 0.0 for j in range(n)] for i in range(n - 1)]
    if flip:  # the other orientation of the same plane: its normal points the other way
        edges = [edges[1], edges[0]] + edges[2:]
    return dict(kind=1, org=org, edges=edges, rgb=list(rgb), diffuse=diffuse, light_rgb=light_rgb)


def scene_cfg(n, shapes, lights, background=(0.2, 0.4, 0.6), background_is_light=True):
    return dict(n=n, camera_org=[0.13 * (i + 1) for i in range(n - 1)] + [CAM_Z],
                camera_dir=[0.0] * (n - 1) + [-1.0],
                axes=[[1.0 if j == i else 0.0 for j in range(n)] for i in range(n - 1)],
                units_per_pixel=UPP, screen=SCREENS[n], projector=0, lights=lights,
                background_rgb=list(background), background_is_light=background_is_light,
                object_rgb=[1.0, 1.0, 1.0], diffuse=0.8, shapes=shapes)


def hit_points(cfg, scene, pixels, S, z):
    """Hit points on the plane x_{n-1} = z of every (pixel, sample) ray:
    ParallelProjector::ray (projector.h:194-201) at screen point
    -size/2 + pixel + sample point, moved along -e_{n-1}."""
    n = cfg["n"]
    scr = np.array(cfg["screen"], np.float64)
    pts = np.zeros((len(pixels), S, n))
    for a, p in enumerate(pixels):
        for s in range(S):
            sp = O.sample_point(scene, int(p), s) - scr / 2
            pts[a, s, :n - 1] = np.array(cfg["camera_org"][:n - 1]) + UPP * sp
            pts[a, s, n - 1] = z
    return pts


def direct_closed_form(n, p, light, normal):
    """One light's term of direct_diffuse_lighting at points p [.., n] with unit
    normal facing the camera, unoccluded: colour [.., 3] float32, and whether
    the light counts (a shadow ray is cast)."""
    vl = np.asarray(light["pos"], np.float64) - p
    r = np.linalg.norm(vl, axis=-1)
    if light.get("kind", 0) == 1:
        scale = np.ones_like(r)
    else:
        u = light["unit_distance"]
        scale = u ** (n - 1) / r ** (n - 1)
    lc = np.asarray(light["rgb"], np.float32)[None] * scale.astype(np.float32)[..., None]
    cos = (vl @ normal) / r
    lit = (lc.max(axis=-1) >= f32(1e-4)) & (cos > 0)
    w = (COEF[n] * cos).astype(np.float32)
    return np.where(lit[..., None], lc * w[..., None], np.float32(0)), lit


def run(n, cfg, depth, pixels=None, seed=3, shift=0.0):
    v, f = far_mesh(n, shift=shift)
    m = O.Mesh(v, f)
    sc = O.make_scene(cfg, 4, depth, seed)
    if pixels is None:
        pixels = np.arange(int(np.prod(cfg["screen"])))
    img, rays = O.render_pixels(m, sc, pixels, threads=2)
    S = O.samples_per_dimension(n - 1, 4) ** (n - 1)
    return img, rays, sc, pixels, S


@pytest.mark.parametrize("n", [3, 4, 5, 6])
def test_stratified_sample_points_lie_in_their_cells(n):
    # engine.h:86-156: k^(n-1) samples, sample s in the cell of its base-k digits
    # (dimension 0 least significant), jitter uniform in the cell
    cfg = scene_cfg(n, [], [])
    for spp, k in ((4, O.samples_per_dimension(n - 1, 4)), (27, O.samples_per_dimension(n - 1, 27))):
        sc = O.make_scene(cfg, spp, 0, 11)
        scr = cfg["screen"]
        P = int(np.prod(scr))
        fr = []
        for pixel in range(0, P, max(1, P // 17)):
            coords, rem = [], pixel
            for size in scr:
                coords.append(rem % size)
                rem //= size
            for s in range(k ** (n - 1)):
                pt = O.sample_point(sc, pixel, s)
                digits = [(s // k ** i) % k for i in range(n - 1)]
                cell = np.floor((pt - np.array(coords)) * k).astype(int)
                assert np.all(np.floor(pt).astype(int) == coords), (pixel, s)
                assert cell.tolist() == digits, (pixel, s, pt)
                fr.append((pt - np.array(coords)) * k - np.array(digits))
        fr = np.concatenate(fr)
        assert abs(fr.mean() - 0.5) < 0.05 and fr.min() >= 0 and fr.max() < 1
    # another pass index gives other jitters in the same cells
    a = O.sample_point(O.make_scene(cfg, 4, 0, 11, pass_index=0), 3, 1)
    b = O.sample_point(O.make_scene(cfg, 4, 0, 11, pass_index=1), 3, 1)
    assert not np.array_equal(a, b) and np.array_equal(np.floor(a * 2), np.floor(b * 2))


@pytest.mark.parametrize("n", [3, 4, 5, 6])
@pytest.mark.parametrize("kind", [0, 1])
def test_direct_light_closed_form(n, kind):
    # depth 0: colour = diffuse * rgb * light colour * coef(N) * cos, light colour
    # rgb u^(N-1)/r^(N-1) (point) or rgb (constant); 1 primary + 2 shadow-ray tests
    # (mesh, then the plane) per sample
    light = dict(kind=kind, pos=[0.7, -0.4, 0.5, 0.3, -0.2][:n - 1] + [3.0], rgb=[1.0, 0.8, 0.5],
                 unit_distance=2.5)
    pl = plane(n, 0.0)
    cfg = scene_cfg(n, [pl], [light])
    img, rays, sc, pixels, S = run(n, cfg, 0)
    p = hit_points(cfg, sc, pixels, S, 0.0)
    nrm = np.zeros(n)
    nrm[n - 1] = 1.0
    d, lit = direct_closed_form(n, p, light, nrm)
    assert lit.all()
    surf = (np.float32(pl["diffuse"]) * np.asarray(pl["rgb"], np.float32))
    col = surf * d
    expect = np.zeros((len(pixels), 3), np.float32)
    for s in range(S):
        expect += col[:, s]
    assert np.allclose(img, expect, rtol=2e-6, atol=0)
    assert rays == len(pixels) * S * 3
    # the same plane with its normal the other way: the normal is turned towards the ray
    img2, rays2, *_ = run(n, scene_cfg(n, [plane(n, 0.0, flip=True)], [light]), 0)
    assert np.array_equal(img2, img) and rays2 == rays


@pytest.mark.parametrize("n", [3, 4, 5, 6])
@pytest.mark.parametrize("bg_light", [True, False])
def test_bounce_to_background_closed_form(n, bg_light):
    # depth >= 1: every bounce from the plane leaves the scene; it returns the
    # background's light colour (its luminance) when the background is a light
    # source, else the background colour; 1 + 2 + 1 rays per sample
    light = dict(kind=0, pos=[0.3] * (n - 1) + [2.0], rgb=[0.7, 0.7, 0.9], unit_distance=2.0)
    pl = plane(n, 0.0)
    bg = (0.2, 0.4, 0.6)
    cfg = scene_cfg(n, [pl], [light], background=bg, background_is_light=bg_light)
    img, rays, sc, pixels, S = run(n, cfg, 3)
    p = hit_points(cfg, sc, pixels, S, 0.0)
    nrm = np.zeros(n)
    nrm[n - 1] = 1.0
    d, _ = direct_closed_form(n, p, light, nrm)
    if bg_light:
        lum = f32(0.2126) * f32(bg[0]) + f32(0.7152) * f32(bg[1]) + f32(0.0722) * f32(bg[2])
        back = np.full(3, lum, np.float32)
    else:
        back = np.asarray(bg, np.float32)
    col = (np.float32(pl["diffuse"]) * np.asarray(pl["rgb"], np.float32)) * (d + back)
    expect = np.zeros((len(pixels), 3), np.float32)
    for s in range(S):
        expect += col[:, s]
    assert np.allclose(img, expect, rtol=2e-6, atol=0)
    assert rays == len(pixels) * S * 4


@pytest.mark.parametrize("n", [3, 4, 5, 6])
def test_bounce_to_emitter_closed_form(n):
    # the camera between the diffuse plane (z = 0) and a large emitting plane
    # (z = 10): every bounce meets the emitter and returns its light colour;
    # shadow rays test the mesh, the plane and the emitter (the light lies
    # below it): 1 + 3 + 1 rays per sample
    light = dict(kind=0, pos=[-0.2] * (n - 1) + [3.0], rgb=[1.0, 1.0, 1.0], unit_distance=3.0)
    pl = plane(n, 0.0)
    em = plane(n, 10.0, half=2e4, rgb=(1.0, 1.0, 1.0), diffuse=1.0, light_rgb=[0.25, 0.5, 2.0])
    cfg = scene_cfg(n, [pl, em], [light])
    img, rays, sc, pixels, S = run(n, cfg, 2, seed=5)
    p = hit_points(cfg, sc, pixels, S, 0.0)
    nrm = np.zeros(n)
    nrm[n - 1] = 1.0
    d, _ = direct_closed_form(n, p, light, nrm)
    col = (np.float32(pl["diffuse"]) * np.asarray(pl["rgb"], np.float32)) * (d + np.asarray(em["light_rgb"],
                                                                                           np.float32))
    expect = np.zeros((len(pixels), 3), np.float32)
    for s in range(S):
        expect += col[:, s]
    assert np.allclose(img, expect, rtol=2e-6, atol=0)
    assert rays == len(pixels) * S * 5


@pytest.mark.parametrize("n", [3, 4, 5, 6])
def test_primary_miss_shows_background_colour(n):
    # a primary ray that meets nothing returns the background colour, not its
    # light colour (painter.cpp:333-340 with diffuse_reflection false): one ray each
    light = dict(kind=0, pos=[0.1] * (n - 1) + [2.0], rgb=[1.0, 1.0, 1.0], unit_distance=2.0)
    bg = (0.2, 0.4, 0.6)
    img, rays, sc, pixels, S = run(n, scene_cfg(n, [], [light], background=bg), 3, shift=100.0)
    expect = np.zeros(3, np.float32)
    for _ in range(S):
        expect += np.asarray(bg, np.float32)
    assert np.array_equal(img, np.tile(expect, (len(pixels), 1)))
    assert rays == len(pixels) * S


@pytest.mark.parametrize("n", [3, 4, 5, 6])
def test_cutoffs_and_unlit_side(n):
    P = int(np.prod(SCREENS[n]))
    S = O.samples_per_dimension(n - 1, 4) ** (n - 1)
    light = dict(kind=0, pos=[0.1] * (n - 1) + [2.0], rgb=[1.0, 1.0, 1.0], unit_distance=2.0)
    # surface colour diffuse * rgb below MIN_COLOR_LEVEL: black, no shadow ray, no bounce
    dark = plane(n, 0.0, rgb=(1e-4, 5e-5, 0.0), diffuse=0.9)
    img, rays, *_ = run(n, scene_cfg(n, [dark], [light]), 3)
    assert np.all(img == 0) and rays == P * S
    # a light whose colour at the point is below MIN_COLOR_LEVEL is skipped (no shadow ray)
    dim = dict(light, unit_distance=1e-3)
    img, rays, *_ = run(n, scene_cfg(n, [plane(n, 0.0)], [dim]), 0)
    assert np.all(img == 0) and rays == P * S
    # a light on the far side of the plane: dot(l, n) <= 0, skipped
    behind = dict(light, pos=[0.1] * (n - 1) + [-2.0])
    img, rays, *_ = run(n, scene_cfg(n, [plane(n, 0.0)], [behind]), 0)
    assert np.all(img == 0) and rays == P * S


def test_single_object_scene_conventions():
    # single_object.h:102-179 with min_screen = max_screen: camera at the centre
    # moved to max[N-1] + |size| on the last axis, looking along -e_{N-1}; axes
    # e_0..e_{N-2}; units per pixel = largest projected extent / max screen; screen
    # edges lround(size_i / max projected * max_screen) clamped to [min, max]
    from math_amd import configs
    rng = np.random.default_rng(4)
    for n in (3, 4, 5, 6):
        v = rng.normal(size=(40, n)).astype(np.float32) * rng.uniform(0.5, 2, size=n).astype(np.float32)
        f = rng.integers(0, 30, size=(25, n)).astype(np.int32)  # vertices 30..39 unused
        used = v[np.unique(f)].astype(np.float64)
        mn, mx = used.min(axis=0), used.max(axis=0)
        size = mx - mn
        centre = mn + size / 2
        maxp = size[:n - 1].max()
        cam = configs.camera_for(v, f, 37)
        expect_cam = centre.copy()
        expect_cam[n - 1] = mx[n - 1] + math.sqrt(float(size @ size))
        assert np.allclose(cam["camera_org"], expect_cam, rtol=0, atol=1e-12)
        assert cam["camera_dir"] == [0.0] * (n - 1) + [-1.0]
        assert np.array_equal(np.array(cam["axes"]), np.eye(n)[:n - 1])
        assert cam["units_per_pixel"] == pytest.approx(maxp / 37, rel=1e-15)
        assert cam["screen"] == [min(max(int(math.floor(size[i] / maxp * 37 + 0.5)), 37), 37) for i in range(n - 1)]
    # the background light is the luminance of the background colour (single_object.h:70-72)
    bg = configs.BACKGROUND_RGB
    assert bg == pytest.approx([((c / 255 + 0.055) / 1.055) ** 2.4 for c in (50, 100, 150)], rel=1e-12)


def bar_order(sizes, h):
    """BarPaintbrush pixel order from the loop nest documented at
    paintbrush.h:33-52 over the reversed screen sizes (:148-164): blocks of h
    along every reversed dimension but the last, the last one by one, then the
    block's pixels; each pixel reversed back, and for 2D screens the second
    coordinate flipped."""
    N = len(sizes)
    r = list(reversed(sizes))
    out = []
    for starts in itertools.product(*[range(0, r[i], h) for i in range(N - 1)]):
        for last in range(r[N - 1]):
            ranges = [range(s, min(r[i], s + h)) for i, s in enumerate(starts)]
            for sub in itertools.product(*ranges):
                p = list(reversed(list(sub) + [last]))
                if N == 2:
                    p[1] = sizes[1] - 1 - p[1]
                out.append(tuple(p))
    return out


@pytest.mark.parametrize("sizes,h", [([5, 3], 2), ([4, 7], 3), ([3, 4, 5], 2), ([2, 3, 2, 3], 2), ([6, 5, 4], 4)])
def test_bar_paintbrush_order(sizes, h):
    exe = os.path.join(ROOT, "tests", "cpp", "test_conventions")
    assert os.path.exists(exe), "built by __graft_entry__.build()"
    r = subprocess.run([exe, "paintbrush", str(h)] + [str(s) for s in sizes], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.split("\n")
    got = [tuple(int(x) for x in ln.split()) for ln in lines[1:] if ln.strip()]
    expect = bar_order(sizes, h)
    assert got == expect
    assert len(set(got)) == int(np.prod(sizes))
    # statistics: the pass count starts at 1; pixels count as they are handed out
    assert lines[0].split() == ["passes", "1", "pixels", str(int(np.prod(sizes)))]
