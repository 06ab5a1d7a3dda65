/*
 * nsp.h -- C ABI of libnsp, the MI355X (gfx950) n-space path tracer.
 *
 * This is the drop-in boundary for the painter path of Math Viewer
 * (templeblock/math, src/painter/).  The reference exposes that path as the
 * C++ template `paint<N,T>()` over virtual scene interfaces; a GPU cannot call
 * those virtuals, so the boundary here is the scene DATA those interfaces
 * carry, passed as plain pointers and sizes.  The C++17 adapter
 * math_amd/csrc/painter/painter.h rebuilds the reference's own API
 * (paint<N,T>, PainterNotifier, PaintObjects, Mesh, VisibleSharedMesh,
 * VisiblePointLight, VisibleParallelProjector, VisibleBarPaintbrush,
 * single_object_scene) on top of these functions.
 *
 * Conventions
 *  - every function returns NSP_OK (0) or a negative NSP_E_* code and never
 *    throws; the message is available from nsp_last_error(ctx);
 *  - all array arguments are host pointers unless the name ends in _device;
 *  - n is the ambient dimension N (3..6, the reference instantiates
 *    paint<N,T> for N in {3,4,5,6}: src/painter/painter.cpp:627-643);
 *  - one context per device; a context and the objects made from it are used
 *    from one host thread at a time.
 */
#ifndef NSP_H
#define NSP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NSP_MAX_DIM 6

#define NSP_OK 0
#define NSP_E_INVALID (-1)  /* bad argument or degenerate input (reference error()) */
#define NSP_E_HIP (-2)      /* HIP runtime error */
#define NSP_E_NOMEM (-3)    /* device allocation failed */
#define NSP_E_STOPPED (-4)  /* stop flag observed */
#define NSP_E_UNSUPPORTED (-5)

typedef struct nsp_context nsp_context;
typedef struct nsp_mesh nsp_mesh;
typedef struct nsp_scene nsp_scene;

/* Library and device setup.  No reference counterpart (the reference is CPU
 * only); every other entry point takes the context. */
int nsp_context_create(int device, nsp_context** out);
void nsp_context_destroy(nsp_context* ctx);
/* Last error message of the calling thread for this context (NULL ctx: for
 * nsp_context_create).  Replaces PainterNotifier::painter_error_message
 * (src/painter/painter.h:37) for the ABI's own failures. */
const char* nsp_last_error(const nsp_context* ctx);
/* Forces the next kernel launch of this context to report failure (host-side
 * fault injection for error-path tests; no kernel is launched). */
int nsp_debug_fail_next_launch(nsp_context* ctx);

/* Mesh<N,T>(const Obj<N>*, Matrix<N+1,N+1,T> vertex_matrix, threads, progress)
 * (src/painter/shapes/mesh.h:66, mesh.cpp:61-135).
 *   vertices       V*n floats, row-major (Obj<N>::vertices(), src/obj/obj.h:72)
 *   facets         F*n vertex indices (Obj<N>::Facet::vertices); facet index =
 *                  position in this array, as in Obj<N>::facets() order
 *   normals        VN*n floats or NULL; facet_normals F*n indices or NULL
 *   vertex_matrix  (n+1)*(n+1) doubles row-major, affine; NULL = identity
 * Builds the FP64 per-simplex hyperplane/barycentric bases and the 2^n-tree on
 * the GPU (tree depth and leaf size of src/painter/shapes/mesh.cpp:33-58). */
int nsp_mesh_create(nsp_context* ctx, int n, const float* vertices, int64_t V, const int32_t* facets, int64_t F,
                    const float* normals, int64_t VN, const int32_t* facet_normals, const double* vertex_matrix,
                    nsp_mesh** out);
void nsp_mesh_destroy(nsp_mesh* mesh);
/* Mesh::min_max (src/painter/shapes/mesh.cpp, used by compute_ray_offset and
 * single_object_scene). */
int nsp_mesh_min_max(const nsp_mesh* mesh, double* min, double* max);

/* OBJ / OBJN files (host only).  nsp_obj_read replaces load_from_obj_file<N>
 * and FileObj<N> (src/obj/file/file_load.cpp:946-1144) with the dimension
 * rule of obj_file_dimension_and_type (src/obj/file/obj_file.cpp:144-199):
 * extension "obj" is n = 3, "objN" is n = N.  Lines v / vt / vn / f are read
 * (vn normalised); a face line of G vertex groups becomes G - (n - 1) facets
 * (group 0 with groups i+1 .. i+n-1; at most 5 facets for n = 3, 1 otherwise);
 * indices are 1-based or negative (relative to the items read before the
 * line); zero indices, texture or normal indices given on only some groups and
 * lines with fewer than n groups are errors.  After reading (file_load.cpp:
 * 1402-1430): no facets, an index out of range or a facet bounding box of
 * zero extent along an axis are errors, and for n = 3 facets whose edge
 * vectors are parallel are dropped.  The file is read in `threads` line
 * chunks in parallel (<= 0: the host's threads, chunks of at least 64 KiB).
 * Errors return NSP_E_INVALID with the message in nsp_obj_last_error(). */
typedef struct nsp_obj nsp_obj;
int nsp_obj_read(const char* path, int threads, nsp_obj** out);
void nsp_obj_destroy(nsp_obj* obj);
/* Last error message of nsp_obj_read / nsp_obj_write on the calling thread. */
const char* nsp_obj_last_error(void);
/* Counts of the file's items (Obj<N>::vertices(), facets(), normals(),
 * texcoords(), src/obj/obj.h:72-80); any pointer may be NULL. */
int nsp_obj_info(const nsp_obj* obj, int* n, int64_t* vertices, int64_t* facets, int64_t* normals,
                 int64_t* texcoords);
/* Copies the arrays in the layout nsp_mesh_create takes (0-based indices;
 * facet_normals -1 where a facet has no normals); any pointer may be NULL. */
int nsp_obj_copy(const nsp_obj* obj, float* vertices, int32_t* facets, float* normals, int32_t* facet_normals);
/* nsp_mesh_create on the arrays of a read file (Mesh<N,T>(obj, matrix, ...),
 * src/painter/shapes/mesh.h:66). */
int nsp_mesh_create_obj(nsp_context* ctx, const nsp_obj* obj, const double* vertex_matrix, nsp_mesh** out);
/* save_obj_geometry_to_file (src/obj/file/file_save.cpp:36-306): an optional
 * "# " comment (NULL: none), all vertices scaled into [-1, 1] about the centre
 * of the bounding box of the facets' vertices (proportions kept), "%12.9f"
 * coordinates, the normals normalised, facets "f a b c", or
 * "f a//na b//nb c//nc" when all of a facet's normal indices are >= 0 (for
 * n = 3 vertices 1 and 2 swap places when all three vertex normals point
 * against the facet's edge cross product).  The extension must be "obj"
 * (n = 3) or "objN" (N = n). */
int nsp_obj_write(const char* path, int n, const float* vertices, int64_t V, const int32_t* facets, int64_t F,
                  const float* normals, int64_t VN, const int32_t* facet_normals, const char* comment);

/* Convex hull of `count` float points in n = 2..6 dimensions: replaces
 * compute_convex_hull<N> (src/geometry/core/convex_hull.h:70-72,
 * convex_hull.cpp:877-886).  The points are shuffled (std::shuffle with
 * std::mt19937_64(count)), scaled to integers in [0, 2^30 - 1] keeping
 * proportions (llround((p - min) * (2^30 - 1) / max extent)), duplicates
 * dropped (the first in shuffled order is kept), and the hull is built
 * incrementally with exact predicates.  Each facet has n vertices (input point
 * indices) and a unit outward normal (ConvexHullFacet<N>::vertices / ortho).
 * Coplanar points are not hull vertices unless needed, and a non-simplicial
 * face comes out triangulated as the insertion order makes it.  The first
 * `first_points` points (0: 4096) are inserted on the host; then, in stages
 * [K, 4K), [4K, 16K), ..., the GPU computes the conflicts of the stage's points
 * with the hull built so far in one pass, drops the points strictly inside it,
 * and the host inserts the rest.  Errors (NSP_E_INVALID): fewer than
 * n + 1 distinct points, all points equal, no n-simplex among them, non-finite
 * coordinates; NSP_E_UNSUPPORTED for n outside 2..6. */
typedef struct nsp_hull nsp_hull;
int nsp_convex_hull(nsp_context* ctx, int n, const float* points, int64_t count, int64_t first_points,
                    nsp_hull** out);
/* Delaunay simplices of points in n = 2..5 dimensions: replaces
 * compute_delaunay<N> (convex_hull.h:67-69, convex_hull.cpp:864-875): the
 * points scaled to integers in [0, 2^24 - 1] as above, lifted to the
 * paraboloid x_{n+1} = |x|^2, and the facets of the lifted hull whose outward
 * normal has a negative last coordinate kept.  Each simplex has n + 1 vertices
 * and n + 1 unit normals: normal r is that of the face without vertex r,
 * pointing away from vertex r (DelaunaySimplex<N>::ortho).  nsp_hull_copy's
 * `points` receives the integer coordinates the simplices were computed on
 * (zero for dropped duplicates), as compute_delaunay's `points` output. */
int nsp_delaunay(nsp_context* ctx, int n, const float* points, int64_t count, int64_t first_points,
                 nsp_hull** out);
/* Sizes: facets (hull) or simplices (Delaunay), vertices and unit normals per
 * facet, and the number of distinct points; any pointer may be NULL. */
int nsp_hull_info(const nsp_hull* h, int* n, int64_t* facets, int* vertices_per_facet, int* orthos_per_facet,
                  int64_t* unique_points);
/* vertices [facets][vertices_per_facet]; orthos [facets][orthos_per_facet][n];
 * points [count][n] (Delaunay only); order [unique_points]: the input index of
 * each distinct point in insertion order.  Any pointer may be NULL. */
int nsp_hull_copy(const nsp_hull* h, int32_t* vertices, double* orthos, double* points, int32_t* order);
/* Points the GPU pass dropped, facets created in total, and the times in ms of
 * the host insertions before the GPU pass, the GPU pass, the host insertions
 * after it (ms[3]). */
int nsp_hull_stats(const nsp_hull* h, int64_t* dropped, int64_t* facets_created, double* ms);
void nsp_hull_destroy(nsp_hull* h);

/* Cocone and BoundCocone surface reconstruction of `count` float points in
 * n = 2..5 dimensions.  nsp_manifold_create replaces
 * create_manifold_constructor<N> (cocone_only = 0) and
 * create_manifold_constructor_cocone<N> (cocone_only = 1)
 * (src/geometry/cocone/reconstruction.h:50-57, reconstruction.cpp:386-421):
 * the Delaunay simplices of nsp_delaunay, their Voronoi vertices, and per
 * point the positive pole and the cocone flags of its Delaunay facets (with
 * cocone_only = 0 also the height and radius of its Voronoi cell).  Errors
 * (NSP_E_INVALID): fewer than n + 2 points ("Error point count ..."), the
 * Delaunay errors, "Voronoi vertex is not finite", and the reference's per-point
 * errors ("Positive pole vector not finite", "Negative pole vector not
 * found", "Cocone intersection not found, ...", "Voronoi cell radius is not
 * positive"); NSP_E_UNSUPPORTED for n outside 2..5.  The messages are in
 * nsp_last_error(ctx). */
typedef struct nsp_manifold nsp_manifold;
int nsp_manifold_create(nsp_context* ctx, int n, const float* points, int64_t count, int cocone_only,
                        nsp_manifold** out);
/* ManifoldConstructor::cocone (reconstruction.cpp:321-339): the facets whose
 * Voronoi edge meets the cocones of all their vertices, pruned at sharp ridges
 * (prune_facets.h:117-200) and reduced to the manifold reachable from outside
 * (extract_manifold.h:105-131).  *facets receives the facet count; the facets
 * and normals are read with nsp_manifold_result.  Errors (NSP_E_INVALID, the
 * reference's messages): "Cocone facets not found[ after prune| after
 * manifold extraction]. (n-1)-manifold is not reconstructable." */
int nsp_manifold_cocone(nsp_context* ctx, nsp_manifold* m, int64_t* facets);
/* ManifoldConstructor::bound_cocone(rho, alpha) (reconstruction.cpp:341-373):
 * interior points (thin Voronoi cells whose poles agree with their cocone
 * neighbours' within alpha, reconstruction.cpp:72-160), the facets whose
 * interior vertices all have the cocone flag, then pruning and extraction as
 * above.  Errors (NSP_E_INVALID): rho or alpha outside (0, 1) ("Rho must be in
 * the interval (0, 1), but rho = ..."), a constructor created with
 * cocone_only, "Interior vertices not found. ...", "Cocone interior facets not
 * found. ...", and those of nsp_manifold_cocone. */
int nsp_manifold_bound_cocone(nsp_context* ctx, nsp_manifold* m, double rho, double alpha, int64_t* facets);
/* The last reconstruction's result: normals [points][n] (the positive pole of
 * each facet vertex, zero for other points) and facets [facets][n] (sorted
 * point indices; facets in the lexicographic order of their tuples).  Either
 * pointer may be NULL. */
int nsp_manifold_result(const nsp_manifold* m, double* normals, int32_t* facets);
/* Point count, Delaunay objects (simplices) and Delaunay facets; any pointer
 * may be NULL. */
int nsp_manifold_info(const nsp_manifold* m, int* n, int64_t* points, int64_t* delaunay_objects,
                      int64_t* delaunay_facets);
/* ManifoldConstructor::delaunay_objects (reconstruction.cpp:375-384): the
 * simplices' vertex indices, [delaunay_objects][n + 1]. */
int nsp_manifold_objects(const nsp_manifold* m, int32_t* objects);
/* Times in ms (ms[6]): Delaunay, facet grouping on the host, the GPU passes of
 * the constructor, and of the last reconstruction its GPU passes, pruning, and
 * extraction with the output. */
int nsp_manifold_stats(const nsp_manifold* m, double* ms);
void nsp_manifold_destroy(nsp_manifold* m);

typedef struct {
    int n;
    int64_t facets;
    int64_t tree_nodes;
    int64_t tree_leaves;
    int64_t tree_entries;
    int tree_levels;
    double basis_ms;  /* FP64 basis build on the GPU */
    double tree_ms;   /* per-level binning tree build on the GPU */
    double root_lo[NSP_MAX_DIM];
    double root_size[NSP_MAX_DIM];
} nsp_mesh_info;
int nsp_mesh_get_info(const nsp_mesh* mesh, nsp_mesh_info* info);

/* FP64 per-simplex data of MeshHyperplaneSimplex (mesh_hyperplane_simplex.cpp:59-143,
 * hyperplane_geometry.h:83-115): normal F*n, plane normals F*(n-1)*n,
 * plane offsets F*(n-1).  Any pointer may be NULL. */
int nsp_mesh_records(const nsp_mesh* mesh, double* normal, double* plane_n, double* plane_d);
/* Leaf lists of the 2^n-tree as a flat int64 stream: per leaf, level, q[n]
 * (integer cell coordinates at that level), count, then count simplex indices.
 * Returns the number of int64 values (call with out=NULL for the size), or a
 * negative error code. */
int64_t nsp_mesh_tree_dump(const nsp_mesh* mesh, int64_t* out, int64_t cap);

/* Which copy of the FP32 candidate records traversal reads (no counterpart in
 * the reference, whose Mesh keeps one record per simplex object):
 *   NSP_LAYOUT_LEAF_SOA    leaf-ordered structure of arrays, one copy per
 *                          leaf entry (coalesced when a wave shares a leaf)
 *   NSP_LAYOUT_SIMPLEX_AOS simplex-ordered 16-byte-aligned records gathered
 *                          through the leaf entry ids (one copy per simplex)
 * Both hold the same FP32 values, so results are identical.  A new mesh
 * uses NSP_LAYOUT_SIMPLEX_AOS, the faster of the two on incoherent bounce rays. */
#define NSP_LAYOUT_LEAF_SOA 0
#define NSP_LAYOUT_SIMPLEX_AOS 1
int nsp_mesh_set_record_layout(nsp_mesh* mesh, int layout);
int nsp_mesh_get_record_layout(const nsp_mesh* mesh);

/* First hit of a ray batch: Mesh::intersect_approximate + intersect_precise
 * (src/painter/shapes/mesh.cpp:137-166) for each ray; the parity entry point.
 *   org, dir  count*n floats (dir need not be unit: Ray normalises it,
 *             src/com/ray.h:34-36)
 *   facet     count int32: simplex index, -1 on miss
 *   t         count doubles: hit distance along the unit direction
 *   near      count bytes (or NULL): 1 when the FP32 decision for some simplex
 *             on the ray fell inside the near-boundary band and was re-taken
 *             in FP64 */
int nsp_trace_first_hit(nsp_mesh* mesh, const float* org, const float* dir, int64_t count, int32_t* facet, double* t,
                        uint8_t* near);
/* Same with device pointers on the given HIP stream (NULL = the context stream);
 * asynchronous. */
int nsp_trace_first_hit_device(nsp_mesh* mesh, const float* org_device, const float* dir_device, int64_t count,
                               int32_t* facet_device, double* t_device, uint8_t* near_device, void* hip_stream);
/* nsp_trace_first_hit without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code that
 * must not call hipMalloc, tight serving loops, harnesses that forbid it).
 * Same arguments, results and errors as nsp_trace_first_hit (host arrays in
 * and out, same kernel); the rays and hits are staged through a fixed 64 MiB
 * array in device memory, one per device, set up by nsp_context_create and
 * held until the process ends.  At most NSP_NOALLOC_MAX_RAYS rays per call
 * (more: NSP_E_INVALID, nothing done).  Every call on one device shares that
 * device's array: calls that overlap on one GPU (several contexts or threads
 * on the same device) race and must be serialised by the caller. */
#define NSP_NOALLOC_MAX_RAYS (1 << 20)
int nsp_trace_first_hit_noalloc(nsp_mesh* mesh, const float* org, const float* dir, int64_t count, int32_t* facet,
                                double* t, uint8_t* near);

/* Scene description: the data behind PaintObjects<N,T> (src/painter/objects.h:212-223). */
#define NSP_LIGHT_POINT 0    /* PointLight, src/painter/lights/light_source.h:26-57 */
#define NSP_LIGHT_CONSTANT 1 /* ConstantLight, src/painter/lights/light_source.h:59-78 */
typedef struct {
    int kind;
    double position[NSP_MAX_DIM];
    float rgb[3];
    double unit_intensity_distance; /* point light only */
} nsp_light;

#define NSP_PROJECTOR_PARALLEL 0    /* ParallelProjector, src/painter/projectors/projector.h:153-202 */
#define NSP_PROJECTOR_PERSPECTIVE 1 /* PerspectiveProjector, projector.h:102-151 */
#define NSP_PROJECTOR_SPHERICAL 2   /* SphericalProjector, projector.h:204-275 */
typedef struct {
    int kind;
    double camera_org[NSP_MAX_DIM];
    double camera_dir[NSP_MAX_DIM];
    double screen_axes[NSP_MAX_DIM - 1][NSP_MAX_DIM];
    double units_per_pixel;      /* parallel */
    int screen_size[NSP_MAX_DIM - 1];
    double view_angle_degrees;   /* perspective and spherical: width_view_angle_degrees */
} nsp_projector;

/* SurfaceProperties (src/painter/objects.h:27-139): colour, diffuse, light source flag/colour. */
typedef struct {
    float rgb[3];
    float diffuse;
    int light_source;
    float light_rgb[3];
} nsp_surface;

/* PaintObjects: objects(), light_sources(), projector(),
 * default_surface_properties().  One mesh object (the configurations of
 * single_object_scene); `object` holds the VisibleSharedMesh surface
 * properties (colour, diffuse). */
int nsp_scene_create(nsp_context* ctx, int n, nsp_mesh* mesh, const nsp_surface* object, const nsp_light* lights,
                     int light_count, const nsp_projector* projector, const nsp_surface* default_surface,
                     nsp_scene** out);
void nsp_scene_destroy(nsp_scene* scene);

/* Flat-shaded shapes beside the mesh (src/painter/visible_shapes.h:31-117):
 *   NSP_SHAPE_PARALLELOTOPE             VisibleParallelotope: org and n edge
 *                                       vectors (Parallelotope, src/painter/space/parallelotope.h)
 *   NSP_SHAPE_HYPERPLANE_PARALLELOTOPE  VisibleHyperplaneParallelotope: org and
 *                                       n-1 edge vectors (space/hyperplane_parallelotope.h)
 * Each has its own surface properties; a shape that is a light source emits
 * light_rgb to diffuse bounces and shows rgb to primary rays (trace_path,
 * painter.cpp:355-358), as the mesh object does. */
#define NSP_SHAPE_PARALLELOTOPE 0
#define NSP_SHAPE_HYPERPLANE_PARALLELOTOPE 1
#define NSP_MAX_SHAPES 64
typedef struct {
    int kind;
    double org[NSP_MAX_DIM];
    double edges[NSP_MAX_DIM][NSP_MAX_DIM]; /* n (parallelotope) or n-1 edge vectors */
    nsp_surface surface;
} nsp_shape;
/* Sets the scene's shapes (replacing any set before).  The scene's objects()
 * are then the mesh followed by the shapes in this order; the order decides
 * the ray counts of shadow rays, which test the objects one after another
 * until one blocks the light (light_source_is_visible, painter.cpp:172-186).
 * Smooth-normal renders of meshes with vertex normals are not supported
 * together with shapes (NSP_E_UNSUPPORTED from nsp_paint). */
int nsp_scene_set_shapes(nsp_scene* scene, const nsp_shape* shapes, int count);

#define NSP_PRECISION_FP32 0 /* FP32 trace with FP64 near-boundary re-test */
#define NSP_PRECISION_FP64 1 /* FP64 everywhere (reference-quality renders) */
typedef struct {
    int samples_per_pixel; /* requested; rounded up to k^(n-1) (sampler.h:31-52) */
    int max_depth;         /* MAX_RECURSION_LEVEL (painter.cpp:42), reference value 100 */
    int smooth_normal;     /* paint(..., smooth_normal) */
    int precision;         /* NSP_PRECISION_* */
    int passes;            /* passes to run (>= 1), or -1: until *stop (VisibleBarPaintbrush
                              max_pass_count, paintbrush.h:141-144) */
    uint64_t seed;         /* counter-based RNG key */
    int64_t batch_paths;   /* paths per wavefront batch; 0 = the largest power of two up to 2^27
                              (FP32) or 2^26 (FP64) whose workspace fits 48 GiB */
    int use_graph;         /* capture a batch's launch sequence in a hipGraph once per call and
                              replay it for every batch of every pass */
    int aov;               /* also accumulate first-hit albedo, normal and direct light per pixel (denoiser inputs) */
    int primary;           /* NSP_PRIMARY_*: how FP32 renders find the primary rays' hits */
} nsp_paint_params;

/* Primary visibility of FP32 renders.  Both give the nearest accepted hit of
 * every primary ray (the same FP32 test with the FP64 re-test inline):
 *   NSP_PRIMARY_AUTO    rasterise the simplices into the parallel projection
 *                       (every primary ray has the projector's direction);
 *                       perspective and spherical projections walk
 *   NSP_PRIMARY_WALK    walk the tree per ray, as for the other bounces */
#define NSP_PRIMARY_AUTO 0
#define NSP_PRIMARY_WALK 1

/* Delivery of results: PainterNotifier::painter_pixel_after(pixel, color)
 * (src/painter/painter.h:26-38) for the pixels of each wavefront batch as the
 * batch finishes (while the next batch runs), in runs of at most 65536
 * consecutive pixels in GlobalIndex order (x0 fastest); rgb holds the running
 * mean colour of each pixel (Pixels::add_color_and_samples, painter.cpp:57-88).
 * pass_done follows the last batch of each pass (BarPaintbrush::next_pass).
 * Called on the thread that called nsp_paint. */
typedef struct {
    void (*pixels_after)(void* user, int64_t first_pixel, int64_t count, const float* rgb_mean);
    void (*pass_done)(void* user, int pass, double seconds);
    void (*error_message)(void* user, const char* message);
    void* user;
} nsp_callbacks;

/* paint<N,T>(notifier, samples_per_pixel, paint_objects, paintbrush,
 * thread_count, stop, smooth_normal) (src/painter/painter.h:40-42,
 * painter.cpp:599-625).  Blocks until `passes` passes are done or *stop becomes
 * nonzero (polled between wavefront batches, the analogue of the per-pixel
 * poll at painter.cpp:419; a callback may set it).  A stopped paint returns
 * NSP_E_STOPPED with the finished batches kept.  Errors are reported through
 * callbacks->error_message as in paint(), and returned. */
int nsp_paint(nsp_scene* scene, const nsp_paint_params* params, const nsp_callbacks* callbacks,
              const volatile int32_t* stop);

/* Paintbrush::statistics (src/painter/paintbrushes/paintbrush.h:222-232):
 * passes, pixels, rays (reference counting rules), samples, last pass time. */
int nsp_statistics(const nsp_scene* scene, int64_t* pass_count, int64_t* pixel_count, int64_t* ray_count,
                   int64_t* sample_count, double* previous_pass_duration);

/* Current per-pixel colour sums and sample counts (host copy): rgb_sum
 * pixels*3 floats, samples pixels int32 (either may be NULL).  The per-pixel
 * state of Pixels (painter.cpp:57-88); with nsp_image_set it is also the
 * checkpoint of a progressive render. */
int nsp_image_get(const nsp_scene* scene, float* rgb_sum, int32_t* samples);
int nsp_image_set(nsp_scene* scene, const float* rgb_sum, const int32_t* samples);
/* Per-pixel sums of the first-hit albedo (pixels*3), of the first-hit normal
 * facing the camera (pixels*n) and of the radiance the first hit receives
 * directly (pixels*3: background or emitter seen by the primary ray plus the
 * unoccluded first-hit light contributions) over the samples of passes painted
 * with aov = 1 (divide by the sample counts of nsp_image_get).  Any pointer may
 * be NULL.  No reference counterpart: these are the denoiser's auxiliary inputs. */
int nsp_image_aov(const nsp_scene* scene, float* albedo_sum, float* normal_sum, float* direct_sum);
/* Per-pixel sums (pixels*3) of the radiance of each pixel's first sample, one
 * per pass painted with aov = 1: a one-sample-per-pixel image per pass, the
 * denoiser's low-sample input.  No reference counterpart. */
int nsp_image_first_sample(const nsp_scene* scene, float* rgb_sum);
/* Number of pixels of the scene's screen. */
int64_t nsp_scene_pixel_count(const nsp_scene* scene);

/* Per-stage timing of the last pass (milliseconds, HIP events on the render
 * stream): raygen, closest-hit traversal, shade, shadow traversal, accumulate;
 * and the number of launches of each. */
typedef struct {
    double ms_raygen, ms_closest, ms_shade, ms_shadow, ms_accumulate, ms_total;
    int64_t launches_closest, launches_shadow;
    int64_t closest_rays, shadow_rays;
    /* traversal counters of the closest-hit stage and of the shadow stage
     * (filled when nsp_set_counters is on) */
    int64_t node_visits, leaf_visits, simplex_tests, retests;
    int64_t shadow_node_visits, shadow_leaf_visits, shadow_simplex_tests, shadow_retests;
    /* rasterised primary visibility (NSP_PRIMARY_AUTO): its time is not part
     * of ms_closest, its rays are part of closest_rays (primary_rays of them) */
    double ms_raster;
    int64_t launches_raster, primary_rays;
    /* FP32 walks over simplex-ordered records: one walk launch per depth d
     * holds the closest-hit walks of depth d + 1's bounce rays and the shadow
     * walks of depth d (with their FP64 decisions); its time is not part of
     * ms_closest or ms_shadow, its traversal counters are split between the
     * two kinds as above */
    double ms_walk;
    int64_t launches_walk;
    /* shadow rays of the primary hits rasterised from the light (one point
     * light, FP32 walks over simplex-ordered records): their time is
     * ms_shadow, they are part of shadow_rays but not of the walks' shadow
     * traversal counters (filled when nsp_set_counters is on) */
    int64_t shadow_rays_rasterised;
} nsp_stage_stats;
int nsp_last_pass_stats(const nsp_scene* scene, nsp_stage_stats* out);
/* Enables per-ray traversal counters (node visits, leaf visits, simplex
 * tests, FP64 re-tests) for roofline accounting; off by default. */
int nsp_set_counters(nsp_scene* scene, int enable);

/* Batched launch: one FP32 pass of every scene in the batch (scenes of any
 * size and dimension, all of the batch's context).  The scenes go to K lanes
 * (a stream, a side stream for the shadow stages and a workspace each; lane 0
 * is the context's) by longest processing time first over the estimate
 * paths x log2(2 + simplices); each lane paints its scenes costliest first
 * with nsp_paint's wavefront batches (batch_paths, 0: the largest power of two
 * up to 2^27 paths within the lane's workspace budget: 48 GiB for up to two
 * lanes, 96 GiB shared among more), and the lanes run concurrently.  With
 * use_graph the whole batch is captured into one hipGraph on the first call
 * and replayed by later calls with the same parameters, lanes, workspaces and
 * scene constants (the pass indices are written before each launch).  No
 * reference counterpart (the reference paints one scene per paint() call);
 * each scene's image and statistics accumulate exactly as after nsp_paint with
 * the same parameters. */
typedef struct nsp_batch nsp_batch;
int nsp_batch_create(nsp_context* ctx, nsp_scene* const* scenes, int count, nsp_batch** out);
void nsp_batch_destroy(nsp_batch* batch);
int nsp_batch_paint(nsp_batch* batch, const nsp_paint_params* params, const volatile int32_t* stop);
/* Number of lanes (1..8, default 2; at most one per scene is used). */
int nsp_batch_set_lanes(nsp_batch* batch, int lanes);
/* rays of the last batch pass (reference counting rules) and its wall time */
int nsp_batch_statistics(const nsp_batch* batch, int64_t* ray_count, double* last_pass_duration);
/* Chooses batch_paths for a scene by timing one whole pass at each candidate
 * queue size (2^20 up to the first power of two holding every path of the
 * screen, within the 48 GiB workspace budget; after one untimed warm-up pass
 * at the largest), and writes the winner into params->batch_paths.  Images do
 * not depend on the choice; the scene's image, AOV sums and statistics are as
 * before the call (the timed passes write no AOVs). */
int nsp_autotune(nsp_scene* scene, nsp_paint_params* params);
/* Chooses the lane count of a batch (1 to 4, kept in the batch) by timing
 * whole batched passes (each after an untimed pass, which captures the graph
 * with use_graph); every scene's image and statistics are restored
 * afterwards. */
int nsp_batch_autotune(nsp_batch* batch, nsp_paint_params* params);
/* The batch's current lane count. */
int nsp_batch_get_lanes(const nsp_batch* batch);

/* Learned denoiser for 3D screens (n = 4 renders).  No reference counterpart
 * (the reference only renders progressively).  8 layers of 3x3x3
 * convolutions, 32 channels, BF16 weights and activations, FP32 accumulation,
 * ReLU between layers; the output is added to the residual input channels
 * [residual_channel, residual_channel+3) (the low-spp radiance), or to zero
 * when residual_channel < 0.  Volumes are channels-last float arrays
 * [nb][depth][height][width][channels] on the host. */
typedef struct nsp_denoiser nsp_denoiser;
int nsp_denoiser_create(nsp_context* ctx, int in_channels, int residual_channel, uint64_t seed, nsp_denoiser** out);
/* Heads of the last layer (no reference counterpart):
 *   NSP_DENOISER_RESIDUAL  3 outputs added to the residual channels (as nsp_denoiser_create)
 *   NSP_DENOISER_KERNEL    27 outputs per voxel are the logits of a 3x3x3 filter; the
 *                          prediction is the softmax-weighted sum, over the taps inside
 *                          the volume, of the channels [residual_channel,
 *                          residual_channel+3) of the tap voxels (kernel prediction, as in
 *                          Bako et al. 2017, "Kernel-predicting convolutional networks for
 *                          denoising Monte Carlo renderings"); the last layer starts at zero,
 *                          a box filter.  The last layer has 32 outputs (27 used). */
#define NSP_DENOISER_RESIDUAL 0
#define NSP_DENOISER_KERNEL 1
int nsp_denoiser_create_head(nsp_context* ctx, int in_channels, int residual_channel, int head, uint64_t seed,
                             nsp_denoiser** out);
void nsp_denoiser_destroy(nsp_denoiser* d);
/* Parameters: per layer, weights [cout][27][cin] then bias [cout], with the
 * padded shapes cin = 16 (layer 0) or 32, cout = 32 or 16 (last layer). */
int64_t nsp_denoiser_param_count(const nsp_denoiser* d);
int nsp_denoiser_get_params(const nsp_denoiser* d, float* out);
int nsp_denoiser_set_params(nsp_denoiser* d, const float* in);
/* Training checkpoint (no reference counterpart): the parameters, Adam's
 * first and second moment estimates (param_count floats each; any pointer may
 * be NULL) and its step count.  The patches of a training step are drawn by
 * (seed, step count), so training resumed from a checkpoint on another
 * denoiser of the same shape continues exactly as the unbroken run. */
int nsp_denoiser_get_state(const nsp_denoiser* d, float* params, float* adam_m, float* adam_v, int64_t* adam_t);
int nsp_denoiser_set_state(nsp_denoiser* d, const float* params, const float* adam_m, const float* adam_v,
                           int64_t adam_t);
/* The denoised image of a scene painted with aov = 1 (n = 4, a 3D screen),
 * without host copies of the input: the network's 13-channel input is built on
 * the device from the scene's per-pixel sums (1-spp radiance: the first-sample
 * sums over the passes, samples_per_pass samples per pixel and pass; 4-spp
 * radiance; albedo; normal), and the prediction is blended with the 4-spp
 * radiance: blend * net + (1 - blend) * input.  output: pixels*3 floats in
 * pixel order, or NULL to keep the result on the device (nsp_denoiser_output
 * copies it).  The denoiser must have 13 input channels and residual channel 3.
 * No reference counterpart. */
int nsp_denoiser_forward_scene(nsp_denoiser* d, const nsp_scene* scene, int samples_per_pass, float blend,
                               float* output);
/* Copies the prediction of the last forward (voxels*3 floats). */
int nsp_denoiser_output(const nsp_denoiser* d, float* output);
/* Device time (HIP events on the denoiser's stream) of the last
 * nsp_denoiser_forward's eight convolutions, without the copies, and their
 * multiply-add count on the real channels (2 * 27 * cin * cout per voxel and
 * layer).  No reference counterpart. */
int nsp_denoiser_last_forward(const nsp_denoiser* d, double* ms, double* flops);
/* Denoised radiance output[nb*depth*height*width*3]. */
int nsp_denoiser_forward(nsp_denoiser* d, const float* input, int nb, int depth, int height, int width,
                         float* output);
/* Mean squared error against target (voxels*3) and its gradient w.r.t. the
 * parameters (param_count floats, may be NULL); no update. */
int nsp_denoiser_loss_grad(nsp_denoiser* d, const float* input, const float* target, int nb, int depth, int height,
                           int width, float* loss, float* grad);
/* Adam training on random patch^3 crops (batch per step) of `volumes`
 * input/target volumes uploaded once; losses[steps] receives each step's loss
 * (copied once at the end: the steps run without a host wait).  The crops of a
 * step are drawn from (seed, the denoiser's Adam step count). */
int nsp_denoiser_train(nsp_denoiser* d, const float* inputs, const float* targets, int volumes, int depth, int height,
                       int width, int patch, int batch, int steps, float lr, uint64_t seed, float* losses);
/* One BF16 implicit-GEMM 3x3x3 convolution (zero padding): x [V][cin],
 * w [cout][27][cin], bias [cout] or NULL, y [V][cout]; cin, cout in {16, 32}.
 * Kernel-level parity entry point for the denoiser's convolution. */
int nsp_conv3d_bf16(nsp_context* ctx, const float* x, int nb, int depth, int height, int width, int cin,
                    const float* w, const float* bias, int cout, int relu, float* y);

/* Synthetic meshes (the configurations' generators; host only, no GPU).
 * Each returns the vertex and facet counts when the output pointers are NULL. */
/* Torus R, r on a nu x nv grid, minor radius per vertex r + U(-noise, noise). n=3. */
int nsp_meshgen_torus(double R, double r, int nu, int nv, double noise, uint64_t seed, float* vertices, int64_t* V,
                      int32_t* facets, int64_t* F);
/* Boundary of [-1,1]^n, each facet split into g^(n-1) cubes and each cube into
 * (n-1)! Kuhn simplices; then optionally projected onto S^(n-1)
 * (sphere != 0), radially scaled by (1 + N(0, sigma)), and with the simplices
 * whose centroid lies within angle cap_angle of a seeded direction removed
 * (cap_angle <= 0: none). */
int nsp_meshgen_cube_sphere(int n, int g, int sphere, double sigma, double cap_angle, uint64_t seed, float* vertices,
                            int64_t* V, int32_t* facets, int64_t* F);

#ifdef __cplusplus
}
#endif
#endif
