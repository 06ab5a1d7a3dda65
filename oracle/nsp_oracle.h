/*
 * nsp_oracle.h -- CPU restatement (FP64) of Math Viewer's n-space painter path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the GPU path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * The product (libnsp.so) never links, loads or calls it.
 *
 * Every function restates a reference function; the citation (file:line,
 * relative to the templeblock/math tree) is given at each definition in
 * nsp_oracle.c.  Arithmetic follows the reference: FP64 geometry, `dot` as an
 * fma chain (src/com/vec.h:258-267), float RGB colour (src/com/color/color.h).
 * Compiled with -ffp-contract=off so that no other fma is formed: the FP64
 * simplex records then match the GPU basis-build kernel bit for bit.
 *
 * Deliberate, documented deviations from the reference (all parity-neutral):
 *  - random numbers come from Philox4x32-10 keyed by (seed, pixel, sample,
 *    depth, stream) instead of per-thread mt19937_64 seeded from /dev/urandom
 *    (src/painter/painter.cpp:46-47, :460); the hemisphere sample is drawn in
 *    closed form with the same distribution as the rejection sampler of
 *    src/painter/sampling/sphere.h:87-116;
 *  - the tree's box/simplex overlap test is the conservative plane test of
 *    src/painter/space/shape_intersection.h:150-189 without the LP that
 *    follows it (:191-278), i.e. a superset of the reference's leaf lists;
 *    first hits do not depend on it because a hit is accepted only inside the
 *    leaf box (src/painter/space/tree.h:549-553);
 *  - the maximum path depth is a parameter (MAX_RECURSION_LEVEL is a
 *    constexpr 100 in src/painter/painter.cpp:42).
 */
#ifndef NSP_ORACLE_H
#define NSP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NMAX 6
#define ORC_MAX_SHAPES 16
#define ORC_MAX_LIGHTS 8

typedef struct orc_mesh orc_mesh;

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). */
void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* cosine_sphere_coefficient(n), src/painter/coefficient/cosine_sphere.h:97-162 */
double orc_cosine_sphere_coefficient(int n);

/* orthogonal_complement_of_unit_vector, src/geometry/core/complement.h:71-200.
 * unit: n doubles; basis_out: (n-1)*n doubles, row-major. */
void orc_orthogonal_complement(int n, const double* unit, double* basis_out);

/* ortho_nn, src/geometry/core/linear_algebra.h:167-236: vectors (n-1)*n. */
void orc_ortho_nn(int n, const double* vectors, double* out);

/* Mesh<N,double> construction without the tree, src/painter/shapes/mesh.cpp:61-125.
 * vertices V*n floats, facets F*n int32, normals VN*n floats or NULL,
 * facet_normals F*n or NULL, matrix (n+1)^2 row-major or NULL (identity).
 * Returns 0 on success, nonzero with a message in orc_last_error() otherwise. */
int orc_mesh_create(int n, const float* vertices, int64_t V, const int32_t* facets, int64_t F,
                    const float* normals, int64_t VN, const int32_t* facet_normals,
                    const double* matrix, orc_mesh** out);
void orc_mesh_destroy(orc_mesh* m);
const char* orc_last_error(void);

/* Per-simplex FP64 records for bit-comparison with the GPU basis build.
 * normal: F*n; plane_n: F*(n-1)*n; plane_d: F*(n-1); vertex0: F*n. Any may be NULL. */
void orc_mesh_records(const orc_mesh* m, double* normal, double* plane_n, double* plane_d, double* vertex0);
void orc_mesh_min_max(const orc_mesh* m, double* mn, double* mx);

/* SpatialSubdivisionTree::decompose, src/painter/space/tree.h:466-525
 * (orc_tree_build: one thread; orc_tree_build_threads: the subtrees below the
 * top levels built by `threads` worker threads, as the reference's BoxJobs
 * workers do; the same boxes and leaf lists in either case). */
int orc_tree_build(orc_mesh* m, int max_depth, int min_objects);
int orc_tree_build_threads(orc_mesh* m, int max_depth, int min_objects, int threads);
/* The conservative box/simplex overlap test the tree build uses (after
 * orc_tree_build, which sets its guard distance): 1 overlap, 0 none, -1 bad
 * arguments.  org/size: the box's low corner and edge lengths. */
int orc_box_simplex_overlap(const orc_mesh* m, int64_t f, const double* org, const double* size);
/* The tree's root box (low corner and edge lengths, with the guard). */
void orc_tree_root(const orc_mesh* m, double* lo, double* size);
/* node_count, leaf_count (non-empty + empty leaves), total leaf entries, max level */
void orc_tree_info(const orc_mesh* m, int64_t* nodes, int64_t* leaves, int64_t* entries, int* depth);
/* Sorted leaf lists keyed by the leaf's (level, integer cell coordinates): for
 * each leaf, level, coords[n], count, then count simplex indices -- written as
 * a flat int64 stream; returns the number of int64 written (or needed when out is NULL). */
int64_t orc_tree_dump(const orc_mesh* m, int64_t* out, int64_t cap);

/* Nearest hit over all simplices (no tree).  idx=-1 on miss.  near is set when
 * the result is within eps_bary (barycentric units) of changing, i.e. when some
 * simplex passes the inflated test but not the deflated one at or before the
 * nearest hit, or when the two nearest hits are within rel_t of each other. */
void orc_trace_brute(const orc_mesh* m, int64_t count, const float* org, const float* dir,
                     double eps_bary, double rel_t, int32_t* idx, double* t, uint8_t* near);

/* Mesh::intersect_approximate + intersect_precise via the tree (tree.h:527-590). */
void orc_trace_tree(const orc_mesh* m, int64_t count, const float* org, const float* dir,
                    int32_t* idx, double* t);

/* Painter scene (single mesh, one point light, one projector), the data of
 * PaintObjects built by single_object_scene (src/painter/scenes/single_object.h:87-179)
 * with VisiblePointLight instead of the constant light. */
typedef struct {
    int n;
    double cam_org[ORC_NMAX];
    double cam_dir[ORC_NMAX];
    double axes[ORC_NMAX - 1][ORC_NMAX];
    double units_per_pixel;
    int screen[ORC_NMAX - 1];
    double light_pos[ORC_NMAX];
    float light_rgb[3];
    double light_unit_distance;
    float background_rgb[3];  /* default surface colour; it is a light of its luminance */
    float object_rgb[3];
    float diffuse;
    int samples_per_pixel;    /* requested; rounded up to k^(n-1) */
    int max_depth;            /* MAX_RECURSION_LEVEL */
    int smooth_normal;
    uint64_t seed;
    int pass_index;
    int projector;             /* 0 parallel, 1 perspective, 2 spherical */
    double view_angle_degrees; /* perspective and spherical */
    /* objects after the mesh, in order (VisibleParallelotope kind 0,
     * VisibleHyperplaneParallelotope kind 1, src/painter/visible_shapes.h:31-117) */
    int shape_count;
    struct {
        int kind;
        double org[ORC_NMAX];
        double edges[ORC_NMAX][ORC_NMAX];
        float rgb[3];
        float diffuse;
        int light;            /* a light source emitting light_rgb */
        float light_rgb[3];
    } shapes[ORC_MAX_SHAPES];
    int no_light;             /* 1: no light source (light_* ignored) */
    int background_not_light; /* 1: the default surface is not a light source */
    int light_kind;           /* of the first light: 0 PointLight, 1 ConstantLight */
    int extra_light_count;    /* lights after the first, in order */
    struct {
        int kind;
        double pos[ORC_NMAX];
        float rgb[3];
        double unit_distance;
    } extra_lights[ORC_MAX_LIGHTS - 1];
} orc_scene;

/* Projector::ray for a screen point in pixel units (the `point` argument):
 * PerspectiveProjector, ParallelProjector, SphericalProjector
 * (src/painter/projectors/projector.h:102-275); dir is unit (Ray ctor).
 * Returns nonzero on an invalid projector. */
int orc_projector_ray(const orc_scene* scene, const double* point, double* org, double* dir);

/* One shape (kind 0 parallelotope with n edges, 1 hyperplane parallelotope with
 * n-1; edges row-major with row stride ORC_NMAX): the nearest t > 0 along the
 * ray (o, d) and the normal there; returns 0 on a miss. */
int orc_shape_intersect(int n, int kind, const double* org, const double* edges, const double* o, const double* d,
                        double* t, double* normal);

/* StratifiedJitteredSampler::one_dimension_size, src/painter/sampling/sampler.h:31-52 */
int orc_samples_per_dimension(int screen_dims, int spp);

/* paint() restated: for pixels [pixel_begin, pixel_begin+pixel_count) in
 * GlobalIndex order (x0 fastest), writes the per-pixel SUM of sample colours
 * (3 floats each) and returns the ray count by the Paintbrush counting rules.
 * threads: worker threads (std::thread workers in the reference). */
int orc_render(const orc_mesh* m, const orc_scene* scene, int64_t pixel_begin, int64_t pixel_count,
               int threads, float* rgb_sum, int64_t* rays);

/* orc_render over an explicit list of pixel indices (GlobalIndex order);
 * rgb_sum[j] belongs to pixels[j]. */
int orc_render_pixels(const orc_mesh* m, const orc_scene* s, const int64_t* pixels, int64_t count, int threads,
                      float* rgb_sum, int64_t* rays);

/* The screen point (pixel units, the pixel's integer coordinates plus the
 * stratified jittered sample point in [0,1)^(n-1)) of one (pixel, sample), as
 * orc_render generates it (painter.cpp:419-433, engine.h:86-156). */
int orc_sample_point(const orc_scene* s, int64_t pixel, int sample, double* point);

/* One path, for per-path comparison with the GPU: returns the path colour. */
void orc_trace_path_sample(const orc_mesh* m, const orc_scene* scene, int64_t pixel, int sample,
                           float* rgb, int64_t* rays);

/* Direction sampler, closed form of random_cosine_weighted_on_hemisphere for
 * the given uniforms (u[0] radius^2, u[1..] tangent direction). */
void orc_cosine_hemisphere(int n, const double* normal, const double* u, double* dir_out);

/* Cocone (mode 0) or BoundCocone(rho, alpha) (mode 1) surface reconstruction
 * (src/geometry/cocone/reconstruction.cpp) of a Delaunay triangulation:
 * points [count][n] (doubles), simplices [S][n+1], orthos [S][n+1][n] (ortho r:
 * the unit normal of the face without vertex r, pointing away from it).
 * Returns the number of facets, or -1 (orc_last_error).  facets_out (when not
 * NULL) receives up to facets_cap sorted vertex tuples [*][n] in facet order;
 * normals_out [count][n] the positive-pole normal of each used point (zero
 * otherwise).  n = 2..5. */
int64_t orc_cocone(int n, const double* points, int64_t count, const int32_t* simplices, const double* orthos,
                   int64_t S, int mode, double rho, double alpha, int32_t* facets_out, int64_t facets_cap,
                   double* normals_out);
/* compute_voronoi_vertex: the centre of the sphere through a simplex's n + 1 points. */
void orc_voronoi_vertex(int n, const double* points, const int32_t* simplex, double* out);

#ifdef __cplusplus
}
#endif
#endif
