"""Workload configurations of the n-space painter benchmarks (synthetic, seeded).

Scene conventions follow single_object_scene (src/painter/scenes/single_object.h:102-179):
camera on the last axis looking along -e_{n-1}, screen axes e_0..e_{n-2},
units per pixel = largest projected extent / largest screen edge, background a
light source of its luminance.  Colours follow the reference's painter test
(src/painter/test/test_painter.cpp:36: background sRGB (50,100,150)); the object
is white with diffuse albedo 0.8; one point light (VisiblePointLight) whose
unit-intensity distance is its distance from the origin.

  C0  n=3 torus R=1 r=0.35, 200x100 grid (40,000 triangles), radius noise
      +-0.01 (seed 0), light (3,3,3), 512^2, 16 spp, depth 5
  C1  n=4 tesseract boundary, 8 x 24^3 cubes x 6 Kuhn tetrahedra (663,552),
      on S^3 with radial N(0, 0.02) (seed 1), light (2,2,2,2), 128^3, 16 spp, depth 5
  C2  n=5 penteract boundary, 10 x 10^4 cells x 24 Kuhn 4-simplices (2.4M before
      the cap), on S^4 with a cap of angular radius 0.5 removed (seed 2), 48^4, 8 spp, depth 4

projector_dict gives the same scenes through a perspective or spherical
projector; cornell_box builds the reference's Cornell box around a 3-space mesh.
"""
import numpy as np

from . import native


def srgb_to_linear(c):
    """srgb_to_rgb, src/com/color/conversion.cpp:188-205."""
    c = c / 255.0
    if c >= 1:
        return 1.0
    if c >= 0.04045:
        return ((c + 0.055) / 1.055) ** 2.4
    if c > 0:
        return c / 12.92
    return 0.0


BACKGROUND_RGB = [srgb_to_linear(50), srgb_to_linear(100), srgb_to_linear(150)]


def mesh_min_max(vertices, facets):
    used = vertices[np.unique(facets.ravel())].astype(np.float64)
    return used.min(axis=0), used.max(axis=0)


def camera_for(vertices, facets, screen_edge):
    """single_object_scene camera with min_screen = max_screen = screen_edge."""
    n = vertices.shape[1]
    vmin, vmax = mesh_min_max(vertices, facets)
    size = vmax - vmin
    center = vmin + size / 2
    max_proj = float(np.max(size[: n - 1]))
    screen = [int(min(max(round(size[i] / max_proj * screen_edge), screen_edge), screen_edge)) for i in range(n - 1)]
    cam = center.copy()
    cam[n - 1] = vmax[n - 1] + float(np.linalg.norm(size))
    cdir = np.zeros(n)
    cdir[n - 1] = -1
    axes = np.eye(n)[: n - 1]
    return dict(camera_org=cam.tolist(), camera_dir=cdir.tolist(), axes=axes.tolist(),
                units_per_pixel=max_proj / screen_edge, screen=screen)


def scene_dict(vertices, facets, screen_edge, light_pos, diffuse=0.8):
    n = vertices.shape[1]
    cfg = dict(n=n, object_rgb=[1.0, 1.0, 1.0], diffuse=diffuse, background_rgb=BACKGROUND_RGB,
               light_pos=list(light_pos), light_rgb=[1.0, 1.0, 1.0],
               light_unit_distance=float(np.linalg.norm(light_pos)))
    cfg.update(camera_for(vertices, facets, screen_edge))
    return cfg


def projector_dict(vertices, facets, screen_edge, light_pos, kind, view_angle, diffuse=0.8):
    """scene_dict seen through a perspective or spherical projector
    (PerspectiveProjector / SphericalProjector, src/painter/projectors/projector.h:102-275):
    the same camera direction and screen axes, the camera on the last axis at
    the distance where the view angle spans the object's largest projected
    extent plus its half-diagonal."""
    cfg = scene_dict(vertices, facets, screen_edge, light_pos, diffuse)
    vmin, vmax = mesh_min_max(vertices, facets)
    size = vmax - vmin
    center = vmin + size / 2
    n = vertices.shape[1]
    half = float(np.max(size[: n - 1])) / 2
    dist = half / np.tan(np.radians(view_angle) / 2) + float(np.linalg.norm(size)) / 2
    cam = center.copy()
    cam[n - 1] = center[n - 1] + dist
    cfg.update(projector=kind, view_angle=float(view_angle), camera_org=cam.tolist())
    cfg.pop("units_per_pixel")
    return cfg


def cornell_box(vertices, facets, width, height, size=None, object_rgb=(1.0, 1.0, 1.0), diffuse=1.0,
                camera_direction=(0.0, 0.0, -1.0), camera_up=(0.0, 1.0, 0.0)):
    """CornellBoxScene::create_scene (src/painter/scenes/cornell_box.cpp:114-210) for a
    3-space mesh: five walls and a lamp (hyperplane parallelotopes, the lamp a light
    source of colour 50), a box (parallelotope), no light sources, a black default
    surface that is not a light, and the 70-degree perspective projector.  The mesh
    is scaled to `size` about the origin (model_vertex_matrix); size defaults to 1."""
    size = 1.0 if size is None else float(size)
    v = np.asarray(vertices, np.float64)
    used = v[np.unique(np.asarray(facets).ravel())]
    lo, hi = used.min(axis=0), used.max(axis=0)
    centre = (lo + hi) / 2
    scale = size / float(np.max(hi - lo))
    mesh_v = ((v - centre) * scale).astype(np.float32)
    size *= 1.5
    cdir = np.asarray(camera_direction, np.float64)
    up_in = np.asarray(camera_up, np.float64)
    right = np.cross(cdir, up_in)
    right /= np.linalg.norm(right)
    up = up_in / np.linalg.norm(up_in)
    d = cdir / np.linalg.norm(cdir)
    view_point = (1.0 / 6) * size * up - size * d
    lower_left = view_point + 0.5 * size * (d - right - up)
    lower_right = view_point + 0.5 * size * (d + right - up)
    upper_left = view_point + 0.5 * size * (d - right + up)
    white, red, green, magenta = [1.0, 1.0, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [1.0, 0.0, 1.0]
    hp, pp = native.SHAPE_HYPERPLANE_PARALLELOTOPE, native.SHAPE_PARALLELOTOPE

    def shape(kind, org, edges, rgb, light_rgb=None):
        return dict(kind=kind, org=list(map(float, org)), edges=[list(map(float, e)) for e in edges], rgb=rgb,
                    diffuse=1.0, light_rgb=light_rgb)
    upper_center = upper_left - 0.001 * size * up + 0.5 * size * right + 0.5 * size * d
    shapes = [
        shape(hp, upper_center - 0.1 * size * d - 0.1 * size * right, [0.2 * size * right, 0.2 * size * d], white,
              [50.0, 50.0, 50.0]),
        shape(hp, lower_left + size * d, [size * right, size * up], white),
        shape(hp, upper_left, [size * d, size * right], white),
        shape(hp, lower_left, [size * d, size * right], white),
        shape(hp, lower_left, [size * d, size * up], red),
        shape(hp, lower_right, [size * d, size * up], green),
        shape(pp, lower_left + size * (0.7 * d + 0.8 * right + 0.1 * up),
              [0.1 * size * right, 0.8 * size * up, 0.1 * size * d], magenta),
    ]
    cfg = dict(n=3, object_rgb=list(object_rgb), diffuse=diffuse, background_rgb=[0.0, 0.0, 0.0],
               background_is_light=False, light_pos=None, projector=native.PROJECTOR_PERSPECTIVE, view_angle=70.0,
               camera_org=view_point.tolist(), camera_dir=d.tolist(), axes=[right.tolist(), up.tolist()],
               screen=[int(width), int(height)], shapes=shapes)
    return mesh_v, np.asarray(facets, np.int32), cfg


CONFIGS = {
    # name: (generator, generator args, screen edge, spp, depth, light)
    "C0": ("torus", dict(R=1.0, r=0.35, nu=200, nv=100, noise=0.01, seed=0), 512, 16, 5, (3.0, 3.0, 3.0)),
    "C1": ("cube_sphere", dict(n=4, g=24, sphere=True, sigma=0.02, cap_angle=0.0, seed=1), 128, 16, 5,
           (2.0, 2.0, 2.0, 2.0)),
    "C2": ("cube_sphere", dict(n=5, g=10, sphere=True, sigma=0.0, cap_angle=0.5, seed=2), 48, 8, 4,
           (2.0, 2.0, 2.0, 2.0, 2.0)),
}


WORKLOADS = {
    "C0": "C0: n=3 plain path trace, 40,000-triangle noisy torus, 512^2 parallel-projected screen, 16 spp, depth 5, "
          "point light (3,3,3)",
    "C1": "C1: n=4 plain path trace, 663,552-tetrahedron noisy S^3 mesh, 128^3 parallel-projected screen, "
          "16 spp requested (27 samples/pixel), depth 5, point light (2,2,2,2)",
    "C2": "C2: n=5 plain path trace, S^4 penteract-boundary mesh (2.4M 4-simplices before a 0.5 rad cap is removed), "
          "48^4 parallel-projected screen, 8 spp requested (16 samples/pixel), depth 4, point light (2,2,2,2,2)",
}


def c4_scenes(count=256):
    """The batched workload's scenes: (n, g, screen edge, mesh seed) for each of
    `count` scenes drawn with seed 300: n uniform in {3, 4, 5}; the cube-sphere
    generator's cells per edge g (n=3: 8..59, 768..41,772 triangles; n=4: 3..15,
    1,296..162,000 tetrahedra; n=5: 2..5, 3,840..150,000 4-simplices); screen
    edges 64..256 (n=3) or 32..96 (n >= 4); mesh seed 300 + scene index."""
    rng = np.random.default_rng(300)
    out = []
    for i in range(count):
        n = int(rng.integers(3, 6))
        if n == 3:
            g, edge = int(rng.integers(8, 60)), int(rng.integers(64, 257))
        elif n == 4:
            g, edge = int(rng.integers(3, 16)), int(rng.integers(32, 97))
        else:
            g, edge = int(rng.integers(2, 6)), int(rng.integers(32, 97))
        out.append((n, g, edge, 300 + i))
    return out


def c4_scene(n, g, edge, seed):
    """Mesh (vertices, facets) and scene dict of one batched-workload scene: the
    cube-sphere on S^(n-1) with radial N(0, 0.02) noise, point light at (2, .., 2)."""
    v, f = native.meshgen_cube_sphere(n, g, True, 0.02, 0.0, seed)
    return v, f, scene_dict(v, f, edge, tuple([2.0] * n))


def hull_points(n, count, seed, on_sphere=False, zero=False):
    """Points for the convex hull: uniform in the unit n-ball, or on its sphere,
    as generate_random_data (geometry/test/test_convex_hull.cpp:41-73) draws
    them (here with numpy's generator); with `zero`, the last coordinate is 0
    except for one apex point at x_n = 1."""
    rng = np.random.default_rng(seed)
    out = np.zeros((0, n))
    k = n - 1 if zero else n
    while len(out) < count:
        v = np.zeros((2 * (count - len(out)) + 16, n))
        v[:, :k] = rng.uniform(-1, 1, (len(v), k))
        ln = np.linalg.norm(v, axis=1)
        keep = (ln <= 1) & (ln > 0)
        v, ln = v[keep], ln[keep]
        if on_sphere:
            v = v / ln[:, None]
        out = np.vstack([out, v])
    p = out[:count].astype(np.float32)
    if zero:
        p[-1, n - 1] = 1
    return p


def sphere_with_notch(n, count, bound=False, seed=None):
    """Points of the reconstruction tests' "sphere with a notch"
    (generate_points_sphere_with_notch, geometry/objects/points.cpp:212-236),
    drawn with numpy's generator (seeded by the count unless `seed` is given):
    uniform on the unit (n-1)-sphere, points with a positive last coordinate z
    scaled along the last axis by 1 - |z^5| / 2.  With `bound`, only the part
    with z >= -0.3 (points.cpp:152-164, COS_FOR_BOUND).  A point is kept only
    if its coordinates rounded at 1e-5 differ from every earlier point's
    (DiscretePoints, points.cpp:70-128)."""
    rng = np.random.default_rng(count if seed is None else seed)
    out, seen = [], set()
    while len(out) < count:
        v = rng.uniform(-1, 1, (2 * (count - len(out)) + 64, n))
        sq = np.sum(v * v, axis=1)
        v = v[(sq <= 1) & (sq > 0)]
        v = v / np.linalg.norm(v, axis=1, keepdims=True)
        if bound:
            v = v[v[:, n - 1] >= -0.3]
        z = v[:, n - 1]
        v[:, n - 1] = np.where(z > 0, z * (1 - np.abs(0.5 * z ** 5)), z)
        for p in v:
            key = tuple(np.rint(p * 100000).astype(np.int64))
            if key in seen:
                continue
            seen.add(key)
            out.append(p)
            if len(out) == count:
                break
    return np.array(out).astype(np.float32)


def clone_points(points, shift=3.0):
    """The points and 2^n shifted copies (clone_object,
    geometry/test/test_reconstruction.cpp:118-146): copy k is shifted by +shift
    along axis i where bit i of k is set and by -shift elsewhere."""
    n = points.shape[1]
    parts = [points]
    for k in range(1 << n):
        s = np.array([shift if (k >> i) & 1 else -shift for i in range(n)], np.float32)
        parts.append(points + s)
    return np.concatenate(parts).astype(np.float32)


def sphere_hull_mesh(ctx, n, count, seed=0):
    """simplex_mesh_of_sphere (painter/shapes/test/sphere_mesh.cpp:65-105): the
    convex hull of `count` random points on the unit (n-1)-sphere, as
    (vertices, facets) for a mesh."""
    p = hull_points(n, count, seed, on_sphere=True)
    h = native.convex_hull(ctx, p)
    return p, h.vertices.astype(np.int32)


def make_mesh(gen, args):
    if gen == "torus":
        return native.meshgen_torus(**args)
    return native.meshgen_cube_sphere(**args)


def config(name, **override):
    """Returns (vertices, facets, scene dict, spp, depth) of a configuration.
    override may replace 'screen', 'spp', 'depth' or generator arguments."""
    gen, args, screen, spp, depth, light = CONFIGS[name]
    args = dict(args)
    for k in list(override):
        if k in args:
            args[k] = override.pop(k)
    screen = override.pop("screen", screen)
    spp = override.pop("spp", spp)
    depth = override.pop("depth", depth)
    if override:
        raise KeyError(f"unknown overrides {sorted(override)}")
    v, f = make_mesh(gen, args)
    return v, f, scene_dict(v, f, screen, light), spp, depth
