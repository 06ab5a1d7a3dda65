"""ctypes binding of the CPU checker (libnsp_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnsp_oracle.so")
NMAX = 6


MAX_SHAPES = 16


class ShapeC(C.Structure):
    _fields_ = [("kind", C.c_int), ("org", C.c_double * NMAX), ("edges", (C.c_double * NMAX) * NMAX),
                ("rgb", C.c_float * 3), ("diffuse", C.c_float), ("light", C.c_int), ("light_rgb", C.c_float * 3)]


MAX_LIGHTS = 8


class LightC(C.Structure):
    _fields_ = [("kind", C.c_int), ("pos", C.c_double * NMAX), ("rgb", C.c_float * 3),
                ("unit_distance", C.c_double)]


class Scene(C.Structure):
    _fields_ = [("n", C.c_int), ("cam_org", C.c_double * NMAX), ("cam_dir", C.c_double * NMAX),
                ("axes", (C.c_double * NMAX) * (NMAX - 1)), ("units_per_pixel", C.c_double),
                ("screen", C.c_int * (NMAX - 1)), ("light_pos", C.c_double * NMAX), ("light_rgb", C.c_float * 3),
                ("light_unit_distance", C.c_double), ("background_rgb", C.c_float * 3),
                ("object_rgb", C.c_float * 3), ("diffuse", C.c_float), ("samples_per_pixel", C.c_int),
                ("max_depth", C.c_int), ("smooth_normal", C.c_int), ("seed", C.c_uint64), ("pass_index", C.c_int),
                ("projector", C.c_int), ("view_angle_degrees", C.c_double), ("shape_count", C.c_int),
                ("shapes", ShapeC * MAX_SHAPES), ("no_light", C.c_int), ("background_not_light", C.c_int),
                ("light_kind", C.c_int), ("extra_light_count", C.c_int), ("extra_lights", LightC * (MAX_LIGHTS - 1))]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        fp, dp, ip, i64p = (C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_int32),
                            C.POINTER(C.c_int64))
        u8p = C.POINTER(C.c_uint8)
        u32p = C.POINTER(C.c_uint32)
        L.orc_philox4x32.argtypes = [u32p, u32p, u32p]
        L.orc_philox4x32.restype = None
        L.orc_cosine_sphere_coefficient.argtypes = [C.c_int]
        L.orc_cosine_sphere_coefficient.restype = C.c_double
        L.orc_orthogonal_complement.argtypes = [C.c_int, dp, dp]
        L.orc_orthogonal_complement.restype = None
        L.orc_ortho_nn.argtypes = [C.c_int, dp, dp]
        L.orc_ortho_nn.restype = None
        L.orc_mesh_create.argtypes = [C.c_int, fp, C.c_int64, ip, C.c_int64, fp, C.c_int64, ip, dp, C.POINTER(P)]
        L.orc_mesh_destroy.argtypes = [P]
        L.orc_mesh_destroy.restype = None
        L.orc_last_error.restype = C.c_char_p
        L.orc_mesh_records.argtypes = [P, dp, dp, dp, dp]
        L.orc_mesh_records.restype = None
        L.orc_mesh_min_max.argtypes = [P, dp, dp]
        L.orc_mesh_min_max.restype = None
        L.orc_tree_build.argtypes = [P, C.c_int, C.c_int]
        L.orc_tree_build_threads.argtypes = [P, C.c_int, C.c_int, C.c_int]
        L.orc_box_simplex_overlap.argtypes = [P, C.c_int64, dp, dp]
        L.orc_tree_root.argtypes = [P, dp, dp]
        L.orc_tree_root.restype = None
        L.orc_tree_info.argtypes = [P, i64p, i64p, i64p, C.POINTER(C.c_int)]
        L.orc_tree_info.restype = None
        L.orc_tree_dump.argtypes = [P, i64p, C.c_int64]
        L.orc_tree_dump.restype = C.c_int64
        L.orc_trace_brute.argtypes = [P, C.c_int64, fp, fp, C.c_double, C.c_double, ip, dp, u8p]
        L.orc_trace_brute.restype = None
        L.orc_trace_tree.argtypes = [P, C.c_int64, fp, fp, ip, dp]
        L.orc_trace_tree.restype = None
        L.orc_samples_per_dimension.argtypes = [C.c_int, C.c_int]
        L.orc_render.argtypes = [P, C.POINTER(Scene), C.c_int64, C.c_int64, C.c_int, fp, i64p]
        L.orc_render_pixels.argtypes = [P, C.POINTER(Scene), i64p, C.c_int64, C.c_int, fp, i64p]
        L.orc_sample_point.argtypes = [C.POINTER(Scene), C.c_int64, C.c_int, dp]
        L.orc_trace_path_sample.argtypes = [P, C.POINTER(Scene), C.c_int64, C.c_int, fp, i64p]
        L.orc_trace_path_sample.restype = None
        L.orc_cosine_hemisphere.argtypes = [C.c_int, dp, dp, dp]
        L.orc_cosine_hemisphere.restype = None
        L.orc_cocone.argtypes = [C.c_int, dp, C.c_int64, ip, dp, C.c_int64, C.c_int, C.c_double, C.c_double, ip,
                                 C.c_int64, dp]
        L.orc_cocone.restype = C.c_int64
        L.orc_voronoi_vertex.argtypes = [C.c_int, dp, ip, dp]
        L.orc_voronoi_vertex.restype = None
        _lib = L
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32(c, k, o)
    return list(o)


def cosine_sphere_coefficient(n):
    return lib().orc_cosine_sphere_coefficient(n)


def orthogonal_complement(u):
    u = np.ascontiguousarray(u, np.float64)
    n = u.shape[0]
    out = np.zeros((n - 1, n))
    lib().orc_orthogonal_complement(n, _p(u, C.c_double), _p(out, C.c_double))
    return out


def ortho_nn(vectors):
    v = np.ascontiguousarray(vectors, np.float64)
    n = v.shape[1]
    out = np.zeros(n)
    lib().orc_ortho_nn(n, _p(v, C.c_double), _p(out, C.c_double))
    return out


def cosine_hemisphere(normal, u):
    nrm = np.ascontiguousarray(normal, np.float64)
    uu = np.zeros(8)
    uu[: len(u)] = u
    out = np.zeros(nrm.shape[0])
    lib().orc_cosine_hemisphere(nrm.shape[0], _p(nrm, C.c_double), _p(uu, C.c_double), _p(out, C.c_double))
    return out


def samples_per_dimension(dims, spp):
    return lib().orc_samples_per_dimension(dims, spp)


class Mesh:
    def __init__(self, vertices, facets, normals=None, facet_normals=None, matrix=None):
        L = lib()
        v = np.ascontiguousarray(vertices, np.float32)
        f = np.ascontiguousarray(facets, np.int32)
        nr = None if normals is None else np.ascontiguousarray(normals, np.float32)
        fn = None if facet_normals is None else np.ascontiguousarray(facet_normals, np.int32)
        mt = None if matrix is None else np.ascontiguousarray(matrix, np.float64)
        h = C.c_void_p()
        rc = L.orc_mesh_create(v.shape[1], _p(v, C.c_float), v.shape[0], _p(f, C.c_int32), f.shape[0],
                               _p(nr, C.c_float), 0 if nr is None else nr.shape[0], _p(fn, C.c_int32),
                               _p(mt, C.c_double), C.byref(h))
        if rc:
            raise ValueError(L.orc_last_error().decode())
        self.h = h
        self.n = v.shape[1]
        self.F = f.shape[0]

    def __del__(self):
        try:
            if self.h:
                lib().orc_mesh_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def records(self):
        n, F = self.n, self.F
        nrm, pn, pd, v0 = np.zeros((F, n)), np.zeros((F, n - 1, n)), np.zeros((F, n - 1)), np.zeros((F, n))
        lib().orc_mesh_records(self.h, _p(nrm, C.c_double), _p(pn, C.c_double), _p(pd, C.c_double),
                               _p(v0, C.c_double))
        return nrm, pn, pd, v0

    def min_max(self):
        mn, mx = np.zeros(self.n), np.zeros(self.n)
        lib().orc_mesh_min_max(self.h, _p(mn, C.c_double), _p(mx, C.c_double))
        return mn, mx

    def build_tree(self, max_depth=None, min_objects=10, threads=None):
        """threads: worker threads of the build (default: the CPUs this process may run on)."""
        if max_depth is None:
            max_depth = {3: 10, 4: 8, 5: 6, 6: 5}[self.n]
        if threads is None:
            threads = len(os.sched_getaffinity(0))
        rc = lib().orc_tree_build_threads(self.h, max_depth, min_objects, threads)
        if rc:
            raise ValueError(lib().orc_last_error().decode())

    def box_overlap(self, f, org, size):
        """The tree build's conservative overlap test of simplex f with a box."""
        o = np.ascontiguousarray(org, np.float64)
        z = np.ascontiguousarray(size, np.float64)
        return lib().orc_box_simplex_overlap(self.h, int(f), _p(o, C.c_double), _p(z, C.c_double))

    def tree_root(self):
        lo, size = np.zeros(self.n), np.zeros(self.n)
        lib().orc_tree_root(self.h, _p(lo, C.c_double), _p(size, C.c_double))
        return lo, size

    def tree_info(self):
        a, b, c, d = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int()
        lib().orc_tree_info(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d))
        return dict(nodes=a.value, leaves=b.value, entries=c.value, depth=d.value)

    def tree_dump(self):
        size = lib().orc_tree_dump(self.h, None, 0)
        out = np.zeros(size, np.int64)
        lib().orc_tree_dump(self.h, _p(out, C.c_int64), size)
        return out

    def trace_brute(self, org, dir, eps_bary=1e-9, rel_t=1e-9):
        o = np.ascontiguousarray(org, np.float32)
        d = np.ascontiguousarray(dir, np.float32)
        k = o.shape[0]
        idx, t, near = np.zeros(k, np.int32), np.zeros(k), np.zeros(k, np.uint8)
        lib().orc_trace_brute(self.h, k, _p(o, C.c_float), _p(d, C.c_float), eps_bary, rel_t, _p(idx, C.c_int32),
                              _p(t, C.c_double), _p(near, C.c_uint8))
        return idx, t, near

    def trace_tree(self, org, dir):
        o = np.ascontiguousarray(org, np.float32)
        d = np.ascontiguousarray(dir, np.float32)
        k = o.shape[0]
        idx, t = np.zeros(k, np.int32), np.zeros(k)
        lib().orc_trace_tree(self.h, k, _p(o, C.c_float), _p(d, C.c_float), _p(idx, C.c_int32), _p(t, C.c_double))
        return idx, t


def make_scene(cfg, spp, max_depth, seed, smooth=False, pass_index=0):
    s = Scene()
    n = cfg["n"]
    s.n = n
    s.cam_org[:n] = cfg["camera_org"]
    s.cam_dir[:n] = cfg["camera_dir"]
    for a in range(n - 1):
        s.axes[a][:n] = cfg["axes"][a]
        s.screen[a] = cfg["screen"][a]
    s.units_per_pixel = cfg.get("units_per_pixel", 0.0)
    s.projector = cfg.get("projector", 0)
    s.view_angle_degrees = cfg.get("view_angle", 0.0)
    lights = cfg.get("lights")
    if lights is None:
        lights = [] if cfg.get("light_pos") is None else [
            dict(kind=0, pos=cfg["light_pos"], rgb=cfg["light_rgb"], unit_distance=cfg["light_unit_distance"])]
    if not lights:
        s.no_light = 1
    else:
        s.light_kind = lights[0].get("kind", 0)
        s.light_pos[:n] = lights[0]["pos"]
        s.light_rgb[:] = lights[0]["rgb"]
        s.light_unit_distance = lights[0].get("unit_distance", 1.0)
        s.extra_light_count = len(lights) - 1
        for k, li in enumerate(lights[1:]):
            e = s.extra_lights[k]
            e.kind = li.get("kind", 0)
            e.pos[:n] = li["pos"]
            e.rgb[:] = li["rgb"]
            e.unit_distance = li.get("unit_distance", 1.0)
    s.background_not_light = 0 if cfg.get("background_is_light", True) else 1
    shapes = cfg.get("shapes") or []
    if len(shapes) > MAX_SHAPES:
        raise ValueError("too many shapes for the checker")
    s.shape_count = len(shapes)
    for k, sh in enumerate(shapes):
        d = s.shapes[k]
        d.kind = sh["kind"]
        d.org[:n] = sh["org"]
        for a, e in enumerate(sh["edges"]):
            d.edges[a][:n] = e
        d.rgb[:] = sh["rgb"]
        d.diffuse = sh.get("diffuse", 1.0)
        d.light = 0 if sh.get("light_rgb") is None else 1
        if sh.get("light_rgb") is not None:
            d.light_rgb[:] = sh["light_rgb"]
    s.background_rgb[:] = cfg["background_rgb"]
    s.object_rgb[:] = cfg["object_rgb"]
    s.diffuse = cfg["diffuse"]
    s.samples_per_pixel = spp
    s.max_depth = max_depth
    s.smooth_normal = int(smooth)
    s.seed = seed
    s.pass_index = pass_index
    return s


def render(mesh, scene, pixel_begin=0, pixel_count=None, threads=1):
    """Returns (per-pixel colour sums [count,3] float32, ray count)."""
    L = lib()
    if pixel_count is None:
        pixel_count = int(np.prod([scene.screen[i] for i in range(scene.n - 1)]))
    out = np.zeros((pixel_count, 3), np.float32)
    rays = C.c_int64()
    rc = L.orc_render(mesh.h, C.byref(scene), pixel_begin, pixel_count, threads, _p(out, C.c_float), C.byref(rays))
    if rc:
        raise ValueError(L.orc_last_error().decode())
    return out, rays.value


def render_pixels(mesh, scene, pixels, threads=1):
    """render over an explicit list of pixel indices: (sums [len,3] float32, ray count)."""
    L = lib()
    px = np.ascontiguousarray(pixels, np.int64)
    out = np.zeros((px.shape[0], 3), np.float32)
    rays = C.c_int64()
    rc = L.orc_render_pixels(mesh.h, C.byref(scene), _p(px, C.c_int64), px.shape[0], threads, _p(out, C.c_float),
                             C.byref(rays))
    if rc:
        raise ValueError(L.orc_last_error().decode())
    return out, rays.value


def sample_point(scene, pixel, sample):
    """Screen point (pixel units) of one (pixel, sample): integer pixel coordinates plus the
    stratified jittered point in [0,1)^(n-1)."""
    n = scene.n
    out = np.zeros(n - 1)
    if lib().orc_sample_point(C.byref(scene), pixel, sample, _p(out, C.c_double)):
        raise ValueError(lib().orc_last_error().decode())
    return out


def path_sample(mesh, scene, pixel, sample):
    rgb = np.zeros(3, np.float32)
    rays = C.c_int64()
    lib().orc_trace_path_sample(mesh.h, C.byref(scene), pixel, sample, _p(rgb, C.c_float), C.byref(rays))
    return rgb, rays.value


def projector_ray(cfg, point):
    """Projector::ray at a screen point in pixel units: (org, unit dir)."""
    s = make_scene(cfg, 1, 1, 0)
    n = cfg["n"]
    p = np.ascontiguousarray(point, np.float64)
    org, d = np.zeros(n), np.zeros(n)
    L = lib()
    L.orc_projector_ray.argtypes = [C.POINTER(Scene), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                    C.POINTER(C.c_double)]
    if L.orc_projector_ray(C.byref(s), _p(p, C.c_double), _p(org, C.c_double), _p(d, C.c_double)):
        raise ValueError("invalid projector")
    return org, d


def shape_intersect(kind, org, edges, o, d):
    """One shape's nearest hit t > 0 along (o, d) and the normal there, or None."""
    n = len(org)
    e = np.zeros((NMAX, NMAX))
    for a, v in enumerate(edges):
        e[a, :n] = v
    org_ = np.ascontiguousarray(org, np.float64)
    o_ = np.ascontiguousarray(o, np.float64)
    d_ = np.ascontiguousarray(d, np.float64)
    t = C.c_double()
    nrm = np.zeros(n)
    L = lib()
    L.orc_shape_intersect.argtypes = [C.c_int, C.c_int] + [C.POINTER(C.c_double)] * 4 + \
        [C.POINTER(C.c_double), C.POINTER(C.c_double)]
    if not L.orc_shape_intersect(n, kind, _p(org_, C.c_double), _p(e, C.c_double), _p(o_, C.c_double),
                                 _p(d_, C.c_double), C.byref(t), _p(nrm, C.c_double)):
        return None
    return t.value, nrm


def cocone(points, simplices, orthos, bound=False, rho=0.3, alpha=0.14):
    """Cocone (or BoundCocone(rho, alpha)) of a Delaunay triangulation: points
    [count, n] (the coordinates the triangulation was computed on), simplices
    [S, n+1], orthos [S, n+1, n].  Returns (facets [F, n] sorted vertex tuples in
    facet order, normals [count, n]); raises ValueError with the checker's
    message when the reconstruction fails."""
    pts = np.ascontiguousarray(points, np.float64)
    sim = np.ascontiguousarray(simplices, np.int32)
    ort = np.ascontiguousarray(orthos, np.float64)
    count, n = pts.shape
    S = sim.shape[0]
    cap = max(1, 4 * S * (n + 1))
    facets = np.zeros((cap, n), np.int32)
    normals = np.zeros((count, n))
    k = lib().orc_cocone(n, _p(pts, C.c_double), count, _p(sim, C.c_int32), _p(ort, C.c_double), S, int(bound),
                         rho, alpha, _p(facets, C.c_int32), cap, _p(normals, C.c_double))
    if k < 0:
        raise ValueError(lib().orc_last_error().decode())
    return facets[:k].copy(), normals


def voronoi_vertex(points, simplex):
    pts = np.ascontiguousarray(points, np.float64)
    sim = np.ascontiguousarray(simplex, np.int32)
    out = np.zeros(pts.shape[1])
    lib().orc_voronoi_vertex(pts.shape[1], _p(pts, C.c_double), _p(sim, C.c_int32), _p(out, C.c_double))
    return out
