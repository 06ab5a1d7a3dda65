"""ctypes binding of libnsp.so (include/nsp.h) for tests and benchmarks.

The library must be built (math_amd/build.py); loading fails loudly otherwise.
There is no CPU fallback: every call here runs the HIP path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnsp.so")
MAX_DIM = 6

NSP_OK = 0
NSP_E_INVALID = -1
NSP_E_HIP = -2
NSP_E_NOMEM = -3
NSP_E_STOPPED = -4
NSP_E_UNSUPPORTED = -5

PROJECTOR_PARALLEL = 0
PROJECTOR_PERSPECTIVE = 1
PROJECTOR_SPHERICAL = 2

PRECISION_FP32 = 0
PRECISION_FP64 = 1
LIGHT_POINT = 0
LIGHT_CONSTANT = 1
LAYOUT_LEAF_SOA = 0
LAYOUT_SIMPLEX_AOS = 1
PRIMARY_AUTO = 0
PRIMARY_WALK = 1


class NspError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"nsp error {code}: {msg}")
        self.code = code


class Light(C.Structure):
    _fields_ = [("kind", C.c_int), ("position", C.c_double * MAX_DIM), ("rgb", C.c_float * 3),
                ("unit_intensity_distance", C.c_double)]


class Projector(C.Structure):
    _fields_ = [("kind", C.c_int), ("camera_org", C.c_double * MAX_DIM), ("camera_dir", C.c_double * MAX_DIM),
                ("screen_axes", (C.c_double * MAX_DIM) * (MAX_DIM - 1)), ("units_per_pixel", C.c_double),
                ("screen_size", C.c_int * (MAX_DIM - 1)), ("view_angle_degrees", C.c_double)]


class Surface(C.Structure):
    _fields_ = [("rgb", C.c_float * 3), ("diffuse", C.c_float), ("light_source", C.c_int),
                ("light_rgb", C.c_float * 3)]


SHAPE_PARALLELOTOPE = 0
SHAPE_HYPERPLANE_PARALLELOTOPE = 1


class Shape(C.Structure):
    _fields_ = [("kind", C.c_int), ("org", C.c_double * MAX_DIM), ("edges", (C.c_double * MAX_DIM) * MAX_DIM),
                ("surface", Surface)]


def make_shape(kind, org, edges, rgb, diffuse=1.0, light_rgb=None):
    """nsp_shape: org and edge vectors; light_rgb makes it a light source."""
    sh = Shape()
    sh.kind = kind
    sh.org[:len(org)] = list(org)
    for a, e in enumerate(edges):
        sh.edges[a][:len(e)] = list(e)
    sh.surface.rgb[:] = list(rgb)
    sh.surface.diffuse = diffuse
    sh.surface.light_source = 0 if light_rgb is None else 1
    if light_rgb is not None:
        sh.surface.light_rgb[:] = list(light_rgb)
    return sh


class PaintParams(C.Structure):
    _fields_ = [("samples_per_pixel", C.c_int), ("max_depth", C.c_int), ("smooth_normal", C.c_int),
                ("precision", C.c_int), ("passes", C.c_int), ("seed", C.c_uint64), ("batch_paths", C.c_int64),
                ("use_graph", C.c_int), ("aov", C.c_int), ("primary", C.c_int)]


PIXELS_AFTER = C.CFUNCTYPE(None, C.c_void_p, C.c_int64, C.c_int64, C.POINTER(C.c_float))
PASS_DONE = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_double)
ERROR_MSG = C.CFUNCTYPE(None, C.c_void_p, C.c_char_p)


class Callbacks(C.Structure):
    _fields_ = [("pixels_after", PIXELS_AFTER), ("pass_done", PASS_DONE), ("error_message", ERROR_MSG),
                ("user", C.c_void_p)]


class MeshInfo(C.Structure):
    _fields_ = [("n", C.c_int), ("facets", C.c_int64), ("tree_nodes", C.c_int64), ("tree_leaves", C.c_int64),
                ("tree_entries", C.c_int64), ("tree_levels", C.c_int), ("basis_ms", C.c_double),
                ("tree_ms", C.c_double), ("root_lo", C.c_double * MAX_DIM), ("root_size", C.c_double * MAX_DIM)]


class StageStats(C.Structure):
    _fields_ = [("ms_raygen", C.c_double), ("ms_closest", C.c_double), ("ms_shade", C.c_double),
                ("ms_shadow", C.c_double), ("ms_accumulate", C.c_double), ("ms_total", C.c_double),
                ("launches_closest", C.c_int64), ("launches_shadow", C.c_int64), ("closest_rays", C.c_int64),
                ("shadow_rays", C.c_int64), ("node_visits", C.c_int64), ("leaf_visits", C.c_int64),
                ("simplex_tests", C.c_int64), ("retests", C.c_int64), ("shadow_node_visits", C.c_int64),
                ("shadow_leaf_visits", C.c_int64), ("shadow_simplex_tests", C.c_int64), ("shadow_retests", C.c_int64),
                ("ms_raster", C.c_double), ("launches_raster", C.c_int64), ("primary_rays", C.c_int64),
                ("ms_walk", C.c_double), ("launches_walk", C.c_int64), ("shadow_rays_rasterised", C.c_int64)]


# every symbol of include/nsp.h
EXPORTS = [
    "nsp_context_create", "nsp_context_destroy", "nsp_last_error", "nsp_debug_fail_next_launch",
    "nsp_mesh_create", "nsp_mesh_destroy", "nsp_mesh_min_max", "nsp_mesh_get_info", "nsp_mesh_records",
    "nsp_mesh_tree_dump", "nsp_trace_first_hit", "nsp_trace_first_hit_device", "nsp_trace_first_hit_noalloc",
    "nsp_scene_create",
    "nsp_scene_destroy", "nsp_paint", "nsp_statistics", "nsp_image_get", "nsp_image_set", "nsp_image_aov", "nsp_image_first_sample",
    "nsp_scene_pixel_count", "nsp_last_pass_stats", "nsp_set_counters", "nsp_meshgen_torus",
    "nsp_meshgen_cube_sphere", "nsp_denoiser_create", "nsp_denoiser_create_head", "nsp_denoiser_destroy",
    "nsp_denoiser_param_count", "nsp_denoiser_get_state", "nsp_denoiser_set_state", "nsp_denoiser_last_forward",
    "nsp_denoiser_forward_scene", "nsp_denoiser_output",
    "nsp_denoiser_get_params", "nsp_denoiser_set_params", "nsp_denoiser_forward", "nsp_denoiser_loss_grad",
    "nsp_denoiser_train", "nsp_conv3d_bf16", "nsp_batch_create", "nsp_batch_destroy", "nsp_batch_paint",
    "nsp_batch_statistics", "nsp_batch_set_lanes", "nsp_batch_get_lanes", "nsp_batch_autotune", "nsp_autotune", "nsp_mesh_set_record_layout", "nsp_mesh_get_record_layout",
    "nsp_obj_read", "nsp_obj_destroy", "nsp_obj_last_error", "nsp_obj_info", "nsp_obj_copy", "nsp_mesh_create_obj",
    "nsp_obj_write", "nsp_scene_set_shapes", "nsp_convex_hull", "nsp_delaunay", "nsp_hull_info", "nsp_hull_copy",
    "nsp_hull_stats", "nsp_hull_destroy", "nsp_manifold_create", "nsp_manifold_cocone", "nsp_manifold_bound_cocone",
    "nsp_manifold_result", "nsp_manifold_info", "nsp_manifold_objects", "nsp_manifold_stats", "nsp_manifold_destroy",
]

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run math_amd/build.py (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        fp = C.POINTER(C.c_float)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        i64p = C.POINTER(C.c_int64)
        L.nsp_context_create.argtypes = [C.c_int, C.POINTER(P)]
        L.nsp_context_destroy.argtypes = [P]
        L.nsp_context_destroy.restype = None
        L.nsp_last_error.argtypes = [P]
        L.nsp_last_error.restype = C.c_char_p
        L.nsp_debug_fail_next_launch.argtypes = [P]
        L.nsp_mesh_create.argtypes = [P, C.c_int, fp, C.c_int64, ip, C.c_int64, fp, C.c_int64, ip, dp, C.POINTER(P)]
        L.nsp_mesh_destroy.argtypes = [P]
        L.nsp_mesh_destroy.restype = None
        L.nsp_mesh_min_max.argtypes = [P, dp, dp]
        L.nsp_mesh_get_info.argtypes = [P, C.POINTER(MeshInfo)]
        L.nsp_mesh_records.argtypes = [P, dp, dp, dp]
        L.nsp_mesh_tree_dump.argtypes = [P, i64p, C.c_int64]
        L.nsp_mesh_tree_dump.restype = C.c_int64
        L.nsp_mesh_set_record_layout.argtypes = [P, C.c_int]
        L.nsp_mesh_get_record_layout.argtypes = [P]
        L.nsp_trace_first_hit.argtypes = [P, fp, fp, C.c_int64, ip, dp, C.POINTER(C.c_uint8)]
        L.nsp_trace_first_hit_device.argtypes = [P, P, P, C.c_int64, P, P, P, P]
        L.nsp_trace_first_hit_noalloc.argtypes = [P, fp, fp, C.c_int64, ip, dp, C.POINTER(C.c_uint8)]
        L.nsp_scene_create.argtypes = [P, C.c_int, P, C.POINTER(Surface), C.POINTER(Light), C.c_int,
                                       C.POINTER(Projector), C.POINTER(Surface), C.POINTER(P)]
        L.nsp_scene_destroy.argtypes = [P]
        L.nsp_scene_destroy.restype = None
        L.nsp_paint.argtypes = [P, C.POINTER(PaintParams), C.POINTER(Callbacks), C.POINTER(C.c_int32)]
        L.nsp_statistics.argtypes = [P, i64p, i64p, i64p, i64p, dp]
        L.nsp_image_get.argtypes = [P, fp, ip]
        L.nsp_image_set.argtypes = [P, fp, ip]
        L.nsp_image_aov.argtypes = [P, fp, fp, fp]
        L.nsp_image_first_sample.argtypes = [P, fp]
        L.nsp_scene_pixel_count.argtypes = [P]
        L.nsp_scene_pixel_count.restype = C.c_int64
        L.nsp_last_pass_stats.argtypes = [P, C.POINTER(StageStats)]
        L.nsp_set_counters.argtypes = [P, C.c_int]
        L.nsp_meshgen_torus.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_uint64, fp, i64p,
                                        ip, i64p]
        L.nsp_meshgen_cube_sphere.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_uint64, fp,
                                              i64p, ip, i64p]
        L.nsp_scene_set_shapes.argtypes = [P, C.POINTER(Shape), C.c_int]
        L.nsp_obj_read.argtypes = [C.c_char_p, C.c_int, C.POINTER(P)]
        L.nsp_obj_destroy.argtypes = [P]
        L.nsp_obj_destroy.restype = None
        L.nsp_obj_last_error.argtypes = []
        L.nsp_obj_last_error.restype = C.c_char_p
        L.nsp_obj_info.argtypes = [P, C.POINTER(C.c_int), i64p, i64p, i64p, i64p]
        L.nsp_obj_copy.argtypes = [P, fp, ip, fp, ip]
        L.nsp_mesh_create_obj.argtypes = [P, P, dp, C.POINTER(P)]
        L.nsp_obj_write.argtypes = [C.c_char_p, C.c_int, fp, C.c_int64, ip, C.c_int64, fp, C.c_int64, ip, C.c_char_p]
        L.nsp_convex_hull.argtypes = [P, C.c_int, fp, C.c_int64, C.c_int64, C.POINTER(P)]
        L.nsp_delaunay.argtypes = [P, C.c_int, fp, C.c_int64, C.c_int64, C.POINTER(P)]
        L.nsp_hull_info.argtypes = [P, C.POINTER(C.c_int), i64p, C.POINTER(C.c_int), C.POINTER(C.c_int), i64p]
        L.nsp_hull_copy.argtypes = [P, ip, dp, dp, ip]
        L.nsp_hull_stats.argtypes = [P, i64p, i64p, dp]
        L.nsp_hull_destroy.argtypes = [P]
        L.nsp_hull_destroy.restype = None
        L.nsp_manifold_create.argtypes = [P, C.c_int, fp, C.c_int64, C.c_int, C.POINTER(P)]
        L.nsp_manifold_cocone.argtypes = [P, P, i64p]
        L.nsp_manifold_bound_cocone.argtypes = [P, P, C.c_double, C.c_double, i64p]
        L.nsp_manifold_result.argtypes = [P, dp, ip]
        L.nsp_manifold_info.argtypes = [P, C.POINTER(C.c_int), i64p, i64p, i64p]
        L.nsp_manifold_objects.argtypes = [P, ip]
        L.nsp_manifold_stats.argtypes = [P, dp]
        L.nsp_manifold_destroy.argtypes = [P]
        L.nsp_manifold_destroy.restype = None
        _lib = L
    return _lib


def _ptr(a, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


# ---------------------------------------------------------------------------
# Mesh generators (host only)

def meshgen_torus(R=1.0, r=0.35, nu=200, nv=100, noise=0.01, seed=0):
    L = lib()
    V, F = C.c_int64(), C.c_int64()
    rc = L.nsp_meshgen_torus(R, r, nu, nv, noise, seed, None, C.byref(V), None, C.byref(F))
    if rc:
        raise NspError(rc, "meshgen_torus")
    v = np.zeros((V.value, 3), np.float32)
    f = np.zeros((F.value, 3), np.int32)
    L.nsp_meshgen_torus(R, r, nu, nv, noise, seed, _ptr(v, C.c_float), C.byref(V), _ptr(f, C.c_int32), C.byref(F))
    return v, f


def meshgen_cube_sphere(n, g, sphere=True, sigma=0.0, cap_angle=0.0, seed=0):
    L = lib()
    V, F = C.c_int64(), C.c_int64()
    rc = L.nsp_meshgen_cube_sphere(n, g, int(sphere), sigma, cap_angle, seed, None, C.byref(V), None, C.byref(F))
    if rc:
        raise NspError(rc, "meshgen_cube_sphere")
    v = np.zeros((V.value, n), np.float32)
    f = np.zeros((F.value, n), np.int32)
    rc = L.nsp_meshgen_cube_sphere(n, g, int(sphere), sigma, cap_angle, seed, _ptr(v, C.c_float), C.byref(V),
                                   _ptr(f, C.c_int32), C.byref(F))
    if rc:
        raise NspError(rc, "meshgen_cube_sphere")
    return v, f


# ---------------------------------------------------------------------------
# OBJ / OBJN files (host only)

class ObjFile:
    """A read OBJ / OBJN file: n, vertices [V, n] float32, facets [F, n] int32
    (0-based), normals [VN, n] float32, facet_normals [F, n] int32 (-1: none),
    texcoords (count of vt lines).  The handle feeds Mesh.from_obj."""

    def __init__(self, path, threads=0):
        L = lib()
        h = C.c_void_p()
        rc = L.nsp_obj_read(os.fsencode(path), int(threads), C.byref(h))
        if rc:
            raise NspError(rc, L.nsp_obj_last_error().decode())
        self._L = L
        self.h = h
        n = C.c_int()
        V, F, VN, VT = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        L.nsp_obj_info(h, C.byref(n), C.byref(V), C.byref(F), C.byref(VN), C.byref(VT))
        self.n = n.value
        self.texcoords = VT.value
        self.vertices = np.zeros((V.value, self.n), np.float32)
        self.facets = np.zeros((F.value, self.n), np.int32)
        self.normals = np.zeros((VN.value, self.n), np.float32)
        self.facet_normals = np.zeros((F.value, self.n), np.int32)
        L.nsp_obj_copy(h, _ptr(self.vertices, C.c_float), _ptr(self.facets, C.c_int32),
                       _ptr(self.normals, C.c_float), _ptr(self.facet_normals, C.c_int32))

    def close(self):
        if self.h:
            self._L.nsp_obj_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_obj(path, threads=0):
    return ObjFile(path, threads)


def write_obj(path, vertices, facets, normals=None, facet_normals=None, comment=None):
    L = lib()
    v = np.ascontiguousarray(vertices, np.float32)
    f = np.ascontiguousarray(facets, np.int32)
    nrm = None if normals is None else np.ascontiguousarray(normals, np.float32)
    fn = None if facet_normals is None else np.ascontiguousarray(facet_normals, np.int32)
    rc = L.nsp_obj_write(os.fsencode(path), v.shape[1], _ptr(v, C.c_float), v.shape[0], _ptr(f, C.c_int32),
                         f.shape[0], _ptr(nrm, C.c_float), 0 if nrm is None else nrm.shape[0], _ptr(fn, C.c_int32),
                         None if comment is None else comment.encode())
    if rc:
        raise NspError(rc, L.nsp_obj_last_error().decode())


# ---------------------------------------------------------------------------
# Convex hull and Delaunay simplices (exact integer predicates, GPU conflict pass)

class Hull:
    """Result of convex_hull / delaunay.  vertices [F, k] int32 (input point
    indices; k = n for a hull, n + 1 for Delaunay), orthos [F, n] (hull: unit
    outward normals) or [F, n + 1, n] (Delaunay: normal r of the face without
    vertex r, away from it), order [unique] (input index of each distinct point
    in insertion order), points [count, n] (Delaunay: the integer coordinates
    used), dropped (points the GPU pass found inside), facets_created, ms (host
    before, GPU pass, host after)."""


def _hull(ctx, points, first, delaunay):
    L = ctx._L
    p = np.ascontiguousarray(points, np.float32)
    if p.ndim != 2:
        raise ValueError("points must be [count, n]")
    h = C.c_void_p()
    fn = L.nsp_delaunay if delaunay else L.nsp_convex_hull
    ctx.check(fn(ctx.h, p.shape[1], _ptr(p, C.c_float), p.shape[0], int(first), C.byref(h)))
    try:
        n, per, orth = C.c_int(), C.c_int(), C.c_int()
        F, U = C.c_int64(), C.c_int64()
        L.nsp_hull_info(h, C.byref(n), C.byref(F), C.byref(per), C.byref(orth), C.byref(U))
        r = Hull()
        r.n = n.value
        r.vertices = np.zeros((F.value, per.value), np.int32)
        shape = (F.value, orth.value, n.value) if delaunay else (F.value, n.value)
        r.orthos = np.zeros(shape, np.float64)
        r.order = np.zeros(U.value, np.int32)
        r.points = np.zeros((p.shape[0], n.value), np.float64) if delaunay else None
        L.nsp_hull_copy(h, _ptr(r.vertices, C.c_int32), _ptr(r.orthos, C.c_double), _ptr(r.points, C.c_double),
                        _ptr(r.order, C.c_int32))
        d, fc = C.c_int64(), C.c_int64()
        ms = (C.c_double * 3)()
        L.nsp_hull_stats(h, C.byref(d), C.byref(fc), ms)
        r.dropped, r.facets_created, r.ms = d.value, fc.value, tuple(ms)
        return r
    finally:
        L.nsp_hull_destroy(h)


def convex_hull(ctx, points, first=0):
    """compute_convex_hull (geometry/core/convex_hull.h:70-72) on ctx's GPU."""
    return _hull(ctx, points, first, False)


def delaunay(ctx, points, first=0):
    """compute_delaunay (geometry/core/convex_hull.h:67-69) on ctx's GPU."""
    return _hull(ctx, points, first, True)


# ---------------------------------------------------------------------------
# Cocone and BoundCocone surface reconstruction

class ManifoldConstructor:
    """create_manifold_constructor / create_manifold_constructor_cocone
    (geometry/cocone/reconstruction.h:50-57) on ctx's GPU.  cocone() and
    bound_cocone(rho, alpha) return (normals [count, n] float64, facets [k, n]
    int32) as ManifoldConstructor::cocone / bound_cocone fill them;
    delaunay_objects() the simplices [S, n + 1]."""

    def __init__(self, ctx, points, cocone_only=False):
        self.ctx = ctx
        L = ctx._L
        p = np.ascontiguousarray(points, np.float32)
        if p.ndim != 2:
            raise ValueError("points must be [count, n]")
        h = C.c_void_p()
        ctx.check(L.nsp_manifold_create(ctx.h, p.shape[1], _ptr(p, C.c_float), p.shape[0], int(cocone_only),
                                        C.byref(h)))
        self.h = h
        n, cnt, S, F = C.c_int(), C.c_int64(), C.c_int64(), C.c_int64()
        L.nsp_manifold_info(h, C.byref(n), C.byref(cnt), C.byref(S), C.byref(F))
        self.n, self.count, self.objects, self.delaunay_facets = n.value, cnt.value, S.value, F.value

    def _result(self, k):
        normals = np.zeros((self.count, self.n), np.float64)
        facets = np.zeros((k, self.n), np.int32)
        self.ctx._L.nsp_manifold_result(self.h, _ptr(normals, C.c_double), _ptr(facets, C.c_int32))
        return normals, facets

    def cocone(self):
        k = C.c_int64()
        self.ctx.check(self.ctx._L.nsp_manifold_cocone(self.ctx.h, self.h, C.byref(k)))
        return self._result(k.value)

    def bound_cocone(self, rho, alpha):
        k = C.c_int64()
        self.ctx.check(self.ctx._L.nsp_manifold_bound_cocone(self.ctx.h, self.h, float(rho), float(alpha), C.byref(k)))
        return self._result(k.value)

    def delaunay_objects(self):
        out = np.zeros((self.objects, self.n + 1), np.int32)
        self.ctx._L.nsp_manifold_objects(self.h, _ptr(out, C.c_int32))
        return out

    def stats(self):
        """ms: Delaunay, facet grouping, constructor GPU passes, reconstruction
        GPU passes, pruning, extraction and output."""
        ms = (C.c_double * 6)()
        self.ctx._L.nsp_manifold_stats(self.h, ms)
        return tuple(ms)

    def close(self):
        if self.h:
            self.ctx._L.nsp_manifold_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------

class Context:
    def __init__(self, device=0):
        self._L = lib()
        h = C.c_void_p()
        rc = self._L.nsp_context_create(device, C.byref(h))
        if rc:
            raise NspError(rc, self._L.nsp_last_error(None).decode())
        self.h = h

    def check(self, rc):
        if rc != NSP_OK:
            raise NspError(rc, self._L.nsp_last_error(self.h).decode())
        return rc

    def close(self):
        if self.h:
            self._L.nsp_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def mesh(self, vertices, facets, normals=None, facet_normals=None, matrix=None):
        return Mesh(self, vertices, facets, normals, facet_normals, matrix)


class Mesh:
    @classmethod
    def from_obj(cls, ctx, obj, matrix=None):
        """Mesh of a read OBJ / OBJN file (an ObjFile, or a path to read)."""
        if not isinstance(obj, ObjFile):
            obj = ObjFile(obj)
        mat = None if matrix is None else np.ascontiguousarray(matrix, np.float64)
        h = C.c_void_p()
        ctx.check(ctx._L.nsp_mesh_create_obj(ctx.h, obj.h, _ptr(mat, C.c_double), C.byref(h)))
        self = cls.__new__(cls)
        self.ctx = ctx
        self.n = obj.n
        self.h = h
        self.F = obj.facets.shape[0]
        return self

    def __init__(self, ctx, vertices, facets, normals=None, facet_normals=None, matrix=None):
        self.ctx = ctx
        L = ctx._L
        v = np.ascontiguousarray(vertices, np.float32)
        f = np.ascontiguousarray(facets, np.int32)
        self.n = v.shape[1]
        nrm = None if normals is None else np.ascontiguousarray(normals, np.float32)
        fn = None if facet_normals is None else np.ascontiguousarray(facet_normals, np.int32)
        mat = None if matrix is None else np.ascontiguousarray(matrix, np.float64)
        h = C.c_void_p()
        ctx.check(L.nsp_mesh_create(ctx.h, self.n, _ptr(v, C.c_float), v.shape[0], _ptr(f, C.c_int32), f.shape[0],
                                    _ptr(nrm, C.c_float), 0 if nrm is None else nrm.shape[0], _ptr(fn, C.c_int32),
                                    _ptr(mat, C.c_double), C.byref(h)))
        self.h = h
        self.F = f.shape[0]

    def close(self):
        if self.h:
            self.ctx._L.nsp_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        i = MeshInfo()
        self.ctx.check(self.ctx._L.nsp_mesh_get_info(self.h, C.byref(i)))
        return i

    def min_max(self):
        mn = np.zeros(self.n)
        mx = np.zeros(self.n)
        self.ctx.check(self.ctx._L.nsp_mesh_min_max(self.h, _ptr(mn, C.c_double), _ptr(mx, C.c_double)))
        return mn, mx

    def records(self):
        n, F = self.n, self.F
        nrm = np.zeros((F, n))
        pn = np.zeros((F, n - 1, n))
        pd = np.zeros((F, n - 1))
        self.ctx.check(self.ctx._L.nsp_mesh_records(self.h, _ptr(nrm, C.c_double), _ptr(pn, C.c_double),
                                                    _ptr(pd, C.c_double)))
        return nrm, pn, pd

    @property
    def record_layout(self):
        return self.ctx._L.nsp_mesh_get_record_layout(self.h)

    @record_layout.setter
    def record_layout(self, layout):
        self.ctx.check(self.ctx._L.nsp_mesh_set_record_layout(self.h, int(layout)))

    def tree_dump(self):
        L = self.ctx._L
        size = L.nsp_mesh_tree_dump(self.h, None, 0)
        if size < 0:
            self.ctx.check(int(size))
        out = np.zeros(size, np.int64)
        L.nsp_mesh_tree_dump(self.h, _ptr(out, C.c_int64), size)
        return out

    def first_hit(self, org, dir):
        o = np.ascontiguousarray(org, np.float32)
        d = np.ascontiguousarray(dir, np.float32)
        cnt = o.shape[0]
        idx = np.zeros(cnt, np.int32)
        t = np.zeros(cnt, np.float64)
        near = np.zeros(cnt, np.uint8)
        self.ctx.check(self.ctx._L.nsp_trace_first_hit(self.h, _ptr(o, C.c_float), _ptr(d, C.c_float), cnt,
                                                       _ptr(idx, C.c_int32), _ptr(t, C.c_double),
                                                       _ptr(near, C.c_uint8)))
        return idx, t, near

    def first_hit_noalloc(self, org, dir):
        """first_hit through nsp_trace_first_hit_noalloc (no device allocation;
        at most NOALLOC_MAX_RAYS rays)."""
        o = np.ascontiguousarray(org, np.float32)
        d = np.ascontiguousarray(dir, np.float32)
        cnt = o.shape[0]
        idx = np.zeros(cnt, np.int32)
        t = np.zeros(cnt, np.float64)
        near = np.zeros(cnt, np.uint8)
        self.ctx.check(self.ctx._L.nsp_trace_first_hit_noalloc(self.h, _ptr(o, C.c_float), _ptr(d, C.c_float), cnt,
                                                               _ptr(idx, C.c_int32), _ptr(t, C.c_double),
                                                               _ptr(near, C.c_uint8)))
        return idx, t, near


NOALLOC_MAX_RAYS = 1 << 20


class Scene:
    def __init__(self, ctx, mesh, cfg):
        """cfg: dict from math_amd.configs (n, camera, axes, units_per_pixel, screen, light, colours)."""
        self.ctx = ctx
        self.mesh = mesh
        n = cfg["n"]
        self.n = n
        L = ctx._L
        obj = Surface()
        obj.rgb[:] = cfg["object_rgb"]
        obj.diffuse = cfg["diffuse"]
        obj.light_source = 0
        bg = Surface()
        bg.rgb[:] = cfg["background_rgb"]
        bg.diffuse = 1.0
        bg.light_source = int(cfg.get("background_is_light", True))
        lum = 0.2126 * cfg["background_rgb"][0] + 0.7152 * cfg["background_rgb"][1] + 0.0722 * cfg["background_rgb"][2]
        bg.light_rgb[:] = [lum, lum, lum]
        lights = cfg.get("lights")
        if lights is None:
            lights = [] if cfg.get("light_pos") is None else [
                dict(kind=LIGHT_POINT, pos=cfg["light_pos"], rgb=cfg["light_rgb"],
                     unit_distance=cfg["light_unit_distance"])]
        nlights = len(lights)
        light = (Light * max(1, nlights))()
        for k, li in enumerate(lights):
            light[k].kind = li.get("kind", LIGHT_POINT)
            light[k].position[:n] = li["pos"]
            light[k].rgb[:] = li["rgb"]
            light[k].unit_intensity_distance = li.get("unit_distance", 1.0)
        pr = Projector()
        pr.kind = cfg.get("projector", PROJECTOR_PARALLEL)
        pr.camera_org[:n] = cfg["camera_org"]
        pr.camera_dir[:n] = cfg["camera_dir"]
        for a in range(n - 1):
            pr.screen_axes[a][:n] = cfg["axes"][a]
            pr.screen_size[a] = cfg["screen"][a]
        pr.units_per_pixel = cfg.get("units_per_pixel", 0.0)
        pr.view_angle_degrees = cfg.get("view_angle", 0.0)
        h = C.c_void_p()
        ctx.check(L.nsp_scene_create(ctx.h, n, mesh.h, C.byref(obj), light if nlights else None, nlights,
                                     C.byref(pr), C.byref(bg), C.byref(h)))
        self.h = h
        if cfg.get("shapes"):
            self.set_shapes([sh if isinstance(sh, Shape) else
                             make_shape(sh["kind"], sh["org"], sh["edges"], sh["rgb"], sh.get("diffuse", 1.0),
                                        sh.get("light_rgb")) for sh in cfg["shapes"]])
        self.pixels = L.nsp_scene_pixel_count(h)
        self.screen = list(cfg["screen"])

    def close(self):
        if self.h:
            self.ctx._L.nsp_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def paint(self, spp, max_depth=5, precision=PRECISION_FP32, passes=1, seed=1, batch_paths=0, smooth=False,
              use_graph=False, stop=None, on_pixels=None, aov=False, primary=PRIMARY_AUTO, on_pass=None):
        """passes=-1 paints until the stop flag (a ctypes.c_int32, which a
        callback may set) is nonzero; a stopped paint raises NspError with
        code NSP_E_STOPPED.  on_pixels(first, count, rgb_mean) follows each
        wavefront batch, on_pass(pass_count, seconds) each pass."""
        p = PaintParams(spp, max_depth, int(smooth), precision, passes, seed, batch_paths, int(use_graph), int(aov),
                        int(primary))
        cb = None
        if on_pixels is not None or on_pass is not None:
            def _px(user, first, count, rgb):
                on_pixels(first, count, np.ctypeslib.as_array(rgb, shape=(count * 3,)).reshape(count, 3).copy())

            def _pass(user, k, secs):
                on_pass(k, secs)
            cb = Callbacks(PIXELS_AFTER(_px) if on_pixels else PIXELS_AFTER(0),
                           PASS_DONE(_pass) if on_pass else PASS_DONE(0), ERROR_MSG(0), None)
        stop_flag = C.c_int32(0) if stop is None else stop
        rc = self.ctx._L.nsp_paint(self.h, C.byref(p), C.byref(cb) if cb else None, C.byref(stop_flag))
        self.ctx.check(rc)

    def image(self):
        s = np.zeros((self.pixels, 3), np.float32)
        c = np.zeros(self.pixels, np.int32)
        self.ctx.check(self.ctx._L.nsp_image_get(self.h, _ptr(s, C.c_float), _ptr(c, C.c_int32)))
        return s, c

    def aov(self):
        """Per-pixel sums of first-hit albedo, normal and direct light (see nsp_image_aov)."""
        a = np.zeros((self.pixels, 3), np.float32)
        nrm = np.zeros((self.pixels, self.n), np.float32)
        dl = np.zeros((self.pixels, 3), np.float32)
        self.ctx.check(self.ctx._L.nsp_image_aov(self.h, _ptr(a, C.c_float), _ptr(nrm, C.c_float),
                                                 _ptr(dl, C.c_float)))
        return a, nrm, dl

    def first_sample(self):
        """Per-pixel sums of the first sample's radiance of each pass painted with aov (nsp_image_first_sample)."""
        r = np.zeros((self.pixels, 3), np.float32)
        self.ctx.check(self.ctx._L.nsp_image_first_sample(self.h, _ptr(r, C.c_float)))
        return r

    def set_shapes(self, shapes):
        """nsp_scene_set_shapes: a list of Shape (make_shape), after the mesh in object order."""
        arr = (Shape * max(1, len(shapes)))(*shapes)
        self.ctx.check(self.ctx._L.nsp_scene_set_shapes(self.h, arr, len(shapes)))

    def set_image(self, s, c):
        s = np.ascontiguousarray(s, np.float32)
        c = np.ascontiguousarray(c, np.int32)
        self.ctx.check(self.ctx._L.nsp_image_set(self.h, _ptr(s, C.c_float), _ptr(c, C.c_int32)))

    def statistics(self):
        a, b, c, d = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        e = C.c_double()
        self.ctx.check(self.ctx._L.nsp_statistics(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d), C.byref(e)))
        return dict(passes=a.value, pixels=b.value, rays=c.value, samples=d.value, last_pass_s=e.value)

    def stage_stats(self):
        s = StageStats()
        self.ctx.check(self.ctx._L.nsp_last_pass_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in StageStats._fields_}

    def set_counters(self, on):
        self.ctx.check(self.ctx._L.nsp_set_counters(self.h, int(on)))


# ---------------------------------------------------------------------------
# Denoiser

def _denoiser_lib():
    L = lib()
    if not getattr(L, "_dn_ready", False):
        P = C.c_void_p
        fp = C.POINTER(C.c_float)
        L.nsp_denoiser_create.argtypes = [P, C.c_int, C.c_int, C.c_uint64, C.POINTER(P)]
        L.nsp_denoiser_create_head.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_uint64, C.POINTER(P)]
        L.nsp_denoiser_destroy.argtypes = [P]
        L.nsp_denoiser_destroy.restype = None
        L.nsp_denoiser_param_count.argtypes = [P]
        L.nsp_denoiser_param_count.restype = C.c_int64
        L.nsp_denoiser_get_params.argtypes = [P, fp]
        L.nsp_denoiser_set_params.argtypes = [P, fp]
        L.nsp_denoiser_get_state.argtypes = [P, fp, fp, fp, C.POINTER(C.c_int64)]
        L.nsp_denoiser_last_forward.argtypes = [P, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.nsp_denoiser_forward_scene.argtypes = [P, P, C.c_int, C.c_float, fp]
        L.nsp_denoiser_output.argtypes = [P, fp]
        L.nsp_denoiser_set_state.argtypes = [P, fp, fp, fp, C.c_int64]
        L.nsp_denoiser_forward.argtypes = [P, fp, C.c_int, C.c_int, C.c_int, C.c_int, fp]
        L.nsp_denoiser_loss_grad.argtypes = [P, fp, fp, C.c_int, C.c_int, C.c_int, C.c_int, fp, fp]
        L.nsp_denoiser_train.argtypes = [P, fp, fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                         C.c_float, C.c_uint64, fp]
        L.nsp_conv3d_bf16.argtypes = [P, fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, fp, fp, C.c_int, C.c_int,
                                      fp]
        L._dn_ready = True
    return L


LAYERS = 8
WIDTH = 32


def layer_shapes(head=0):
    """(cout, cin) padded shapes of the denoiser layers (head: HEAD_RESIDUAL or HEAD_KERNEL)."""
    last = 16 if head == 0 else 32
    return [(WIDTH if l < LAYERS - 1 else last, 16 if l == 0 else WIDTH) for l in range(LAYERS)]


def conv3d_bf16(ctx, x, w, bias=None, relu=False):
    """x [nb,d,h,w,cin] f32, w [cout,27,cin] f32 -> y [nb,d,h,w,cout] f32 (GPU, BF16 MFMA)."""
    L = _denoiser_lib()
    x = np.ascontiguousarray(x, np.float32)
    w = np.ascontiguousarray(w, np.float32)
    nb, d, h, wd, cin = x.shape
    cout = w.shape[0]
    y = np.zeros((nb, d, h, wd, cout), np.float32)
    b = None if bias is None else np.ascontiguousarray(bias, np.float32)
    ctx.check(L.nsp_conv3d_bf16(ctx.h, _ptr(x, C.c_float), nb, d, h, wd, cin, _ptr(w, C.c_float),
                                _ptr(b, C.c_float), cout, int(relu), _ptr(y, C.c_float)))
    return y


HEAD_RESIDUAL = 0
HEAD_KERNEL = 1


class Denoiser:
    def __init__(self, ctx, in_channels, residual_channel=-1, seed=0, head=HEAD_RESIDUAL):
        self.ctx = ctx
        L = _denoiser_lib()
        h = C.c_void_p()
        ctx.check(L.nsp_denoiser_create_head(ctx.h, in_channels, residual_channel, head, seed, C.byref(h)))
        self.head = head
        self.h = h
        self.cin = in_channels
        self.nparams = L.nsp_denoiser_param_count(h)

    def close(self):
        if self.h:
            _denoiser_lib().nsp_denoiser_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.zeros(self.nparams, np.float32)
        self.ctx.check(_denoiser_lib().nsp_denoiser_get_params(self.h, _ptr(p, C.c_float)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams
        self.ctx.check(_denoiser_lib().nsp_denoiser_set_params(self.h, _ptr(p, C.c_float)))

    def forward_scene(self, scene, samples_per_pass, blend=1.0, copy=True):
        """Denoised image [pixels, 3] of a scene painted with aov (nsp_denoiser_forward_scene);
        copy=False leaves it on the device (output() copies it)."""
        out = np.zeros((scene.pixels, 3), np.float32) if copy else None
        self.ctx.check(_denoiser_lib().nsp_denoiser_forward_scene(self.h, scene.h, int(samples_per_pass), float(blend),
                                                                  _ptr(out, C.c_float)))
        return out

    def output(self, pixels):
        out = np.zeros((pixels, 3), np.float32)
        self.ctx.check(_denoiser_lib().nsp_denoiser_output(self.h, _ptr(out, C.c_float)))
        return out

    def last_forward(self):
        """(device ms of the last forward's convolutions, their multiply-add FLOPs)."""
        ms, fl = C.c_double(), C.c_double()
        self.ctx.check(_denoiser_lib().nsp_denoiser_last_forward(self.h, C.byref(ms), C.byref(fl)))
        return ms.value, fl.value

    def state(self):
        """(params, Adam first moments, Adam second moments, Adam step count)."""
        p, m, v = (np.zeros(self.nparams, np.float32) for _ in range(3))
        t = C.c_int64()
        self.ctx.check(_denoiser_lib().nsp_denoiser_get_state(self.h, _ptr(p, C.c_float), _ptr(m, C.c_float),
                                                              _ptr(v, C.c_float), C.byref(t)))
        return p, m, v, t.value

    def set_state(self, p, m, v, t):
        p, m, v = (np.ascontiguousarray(a, np.float32) for a in (p, m, v))
        assert p.size == m.size == v.size == self.nparams
        self.ctx.check(_denoiser_lib().nsp_denoiser_set_state(self.h, _ptr(p, C.c_float), _ptr(m, C.c_float),
                                                              _ptr(v, C.c_float), int(t)))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        nb, d, h, w, c = x.shape
        assert c == self.cin
        out = np.zeros((nb, d, h, w, 3), np.float32)
        self.ctx.check(_denoiser_lib().nsp_denoiser_forward(self.h, _ptr(x, C.c_float), nb, d, h, w,
                                                            _ptr(out, C.c_float)))
        return out

    def loss_grad(self, x, target):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        nb, d, h, w, _ = x.shape
        loss = C.c_float()
        g = np.zeros(self.nparams, np.float32)
        self.ctx.check(_denoiser_lib().nsp_denoiser_loss_grad(self.h, _ptr(x, C.c_float), _ptr(t, C.c_float), nb, d,
                                                              h, w, C.byref(loss), _ptr(g, C.c_float)))
        return loss.value, g

    def train(self, inputs, targets, patch=32, batch=16, steps=100, lr=1e-3, seed=0):
        x = np.ascontiguousarray(inputs, np.float32)
        t = np.ascontiguousarray(targets, np.float32)
        S, d, h, w, c = x.shape
        losses = np.zeros(steps, np.float32)
        self.ctx.check(_denoiser_lib().nsp_denoiser_train(self.h, _ptr(x, C.c_float), _ptr(t, C.c_float), S, d, h, w,
                                                          patch, batch, steps, lr, seed, _ptr(losses, C.c_float)))
        return losses


# ---------------------------------------------------------------------------
# Batched persistent launch and autotuner

def _batch_lib():
    L = lib()
    if not getattr(L, "_batch_ready", False):
        P = C.c_void_p
        L.nsp_batch_create.argtypes = [P, C.POINTER(P), C.c_int, C.POINTER(P)]
        L.nsp_batch_destroy.argtypes = [P]
        L.nsp_batch_destroy.restype = None
        L.nsp_batch_paint.argtypes = [P, C.POINTER(PaintParams), C.POINTER(C.c_int32)]
        L.nsp_batch_statistics.argtypes = [P, C.POINTER(C.c_int64), C.POINTER(C.c_double)]
        L.nsp_batch_set_lanes.argtypes = [P, C.c_int]
        L.nsp_batch_get_lanes.argtypes = [P]
        L.nsp_batch_autotune.argtypes = [P, C.POINTER(PaintParams)]
        L.nsp_autotune.argtypes = [P, C.POINTER(PaintParams)]
        L._batch_ready = True
    return L


class Batch:
    def __init__(self, ctx, scenes):
        self.ctx = ctx
        self.scenes = list(scenes)
        L = _batch_lib()
        arr = (C.c_void_p * len(self.scenes))(*[s.h.value for s in self.scenes])
        h = C.c_void_p()
        ctx.check(L.nsp_batch_create(ctx.h, arr, len(self.scenes), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            _batch_lib().nsp_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_lanes(self, lanes):
        self.ctx.check(_batch_lib().nsp_batch_set_lanes(self.h, int(lanes)))

    def paint(self, spp, max_depth=5, seed=1, batch_paths=0, use_graph=False, primary=PRIMARY_AUTO):
        p = PaintParams(spp, max_depth, 0, PRECISION_FP32, 1, seed, batch_paths, int(use_graph), 0, int(primary))
        stop = C.c_int32(0)
        self.ctx.check(_batch_lib().nsp_batch_paint(self.h, C.byref(p), C.byref(stop)))

    def statistics(self):
        r, t = C.c_int64(), C.c_double()
        self.ctx.check(_batch_lib().nsp_batch_statistics(self.h, C.byref(r), C.byref(t)))
        return dict(rays=r.value, last_pass_s=t.value)

    @property
    def lanes(self):
        return _batch_lib().nsp_batch_get_lanes(self.h)

    def autotune(self, spp, max_depth=5, seed=1, use_graph=False):
        """Lane count chosen by timing (kept in the batch and returned); images are restored."""
        p = PaintParams(spp, max_depth, 0, PRECISION_FP32, 1, seed, 0, int(use_graph), 0, 0)
        self.ctx.check(_batch_lib().nsp_batch_autotune(self.h, C.byref(p)))
        return self.lanes


def autotune(scene, spp, max_depth=5, seed=1):
    """Queue size (paths per wavefront batch) chosen by timing; the scene's image is restored."""
    p = PaintParams(spp, max_depth, 0, PRECISION_FP32, 1, seed, 0, 0, 0)
    scene.ctx.check(_batch_lib().nsp_autotune(scene.h, C.byref(p)))
    return p.batch_paths
