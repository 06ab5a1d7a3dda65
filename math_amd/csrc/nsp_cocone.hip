// nsp_cocone.hip -- Cocone and BoundCocone surface reconstruction
// (ManifoldConstructor, src/geometry/cocone/reconstruction.h:38-57,
// reconstruction.cpp:280-408).
//
// The Delaunay simplices come from nsp_delaunay (exact integer predicates).
// The FP64 geometry runs on the GPU: one thread per simplex computes its
// Voronoi vertex; one thread per point computes its positive pole, the height
// and radius of its Voronoi cell and the cocone flags of its facets, and the
// BoundCocone interior-vertex rules read those back per point.  Grouping the
// simplices' faces into facets, pruning the facets at sharp ridges and
// extracting the manifold are sequential walks over small records and run on
// the host, in fixed orders: facets sorted by vertex tuple, a facet's first
// simplex the one with the smaller index, the suspicious ridges of each pruning
// round in sorted order.  (The reference keeps facets and ridges in hash
// containers, whose orders are the standard library's.)
//
// Compiled with -ffp-contract=off: only the explicit fma of the dot products
// and of the Gaussian elimination is fused, as in the reference's dot()
// (com/vec.h) and solve_gauss (numerical/gauss.h:36-100).
#include <hip/hip_runtime.h>

#include <float.h>
#include <math.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "nsp.h"
#include "nsp_device.hpp"
#include "nsp_internal.hpp"

namespace nsp {
// from nsp_api.cpp
hipStream_t context_stream(nsp_context* ctx);
int context_device(nsp_context* ctx);
std::string* context_error(nsp_context* ctx);
}  // namespace nsp

struct nsp_manifold {
    int n = 0;
    int64_t count = 0;
    bool cocone_only = false;
    int device = 0;
    hipStream_t stream = nullptr;
    // the Delaunay triangulation: integer coordinates (input numbering) and simplices
    std::vector<double> points;
    std::vector<int32_t> simplices;  // [S][n + 1]
    int64_t S = 0;
    // facets, sorted by vertex tuple
    int64_t F = 0;
    std::vector<int32_t> fv;        // [F][n] sorted vertex indices
    std::vector<int32_t> fs0, fs1;  // incident simplices (fs1 < 0: one-sided)
    std::vector<double> pn;         // [count][n] positive pole directions
    // device copies kept for the reconstructions
    nsp::DevBuf d_fv, d_fcoc, d_pn, d_height, d_radius, d_vf_start, d_vf_fac, d_vf_loc, d_interior, d_keep, d_flag;
    // result of the last reconstruction
    std::vector<int32_t> out_facets;  // [k][n]
    std::vector<double> out_normals;  // [count][n]
    double ms[6] = {0, 0, 0, 0, 0, 0};
};

namespace nsp {
namespace {

constexpr int CC_BLOCK = 256;

// cos(3 pi / 8): COS_OF_AN_OPENING_ANGLE_WITH_THE_AXIS (cocone.h:29-31)
constexpr double COS_COCONE = 0.38268343236508977172845998403039886676134456248562704143380063562754603396;
// structure.cpp:30-32
constexpr double LIMIT_COS_PA_POLE = 0.99;
constexpr double LIMIT_COS_PA_AB = -0.9999;

// error codes of the per-point pass, in the order of structure.cpp's checks
enum : int8_t {
    VE_NONE = 0,
    VE_POSITIVE_NOT_FINITE,
    VE_NEGATIVE_NOT_FOUND,
    VE_NEGATIVE_NOT_FINITE,
    VE_PA_CLOSE,
    VE_PA_FAR,
    VE_DISTANCE_NOT_FINITE,
    VE_RADIUS,
};

const char* vertex_error_text(int e)
{
    switch (e) {
    case VE_POSITIVE_NOT_FINITE: return "Positive pole vector not finite";
    case VE_NEGATIVE_NOT_FOUND: return "Negative pole vector not found";
    case VE_NEGATIVE_NOT_FINITE: return "Negative pole vector not finite";
    case VE_PA_CLOSE: return "Cocone intersection not found, PA is close to positive pole";
    case VE_PA_FAR: return "Cocone intersection not found, PA is far from positive pole";
    case VE_DISTANCE_NOT_FINITE: return "Cocone intersection distance is not finite";
    case VE_RADIUS: return "Voronoi cell radius is not positive";
    }
    return "unknown error";
}

// dot (com/vec.h): the first product, then fma of the others in order
template <int N>
__host__ __device__ inline double dotn(const double* a, const double* b)
{
    double r = a[0] * b[0];
#pragma unroll
    for (int i = 1; i < N; ++i) r = fma(a[i], b[i], r);
    return r;
}

template <int N>
__host__ __device__ inline void normalizen(double* v)
{
    const double len = sqrt(dotn<N>(v, v));
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = v[i] / len;
}

template <int N>
__host__ __device__ inline bool finiten(const double* v)
{
    bool ok = true;
#pragma unroll
    for (int i = 0; i < N; ++i) ok = ok && isfinite(v[i]);
    return ok;
}

// ---------------------------------------------------------------- Voronoi vertices

// solve_gauss (numerical/gauss.h:36-100): partial pivoting on the largest
// absolute value at or below the diagonal (the first on ties), fma updates.
// Every index is a compile-time constant so that the matrix stays in registers.
template <int N>
__device__ inline void gauss_solve(double (&A)[N][N], double (&b)[N])
{
#pragma unroll
    for (int k = 0; k < N - 1; ++k) {
        double mx = fabs(A[k][k]);
        int pivot = k;
#pragma unroll
        for (int r = k + 1; r < N; ++r) {
            const double v = fabs(A[r][k]);
            if (v > mx) {
                mx = v;
                pivot = r;
            }
        }
#pragma unroll
        for (int r = k + 1; r < N; ++r) {
            if (r == pivot) {
#pragma unroll
                for (int j = 0; j < N; ++j) {
                    const double t = A[r][j];
                    A[r][j] = A[k][j];
                    A[k][j] = t;
                }
                const double t = b[r];
                b[r] = b[k];
                b[k] = t;
            }
        }
#pragma unroll
        for (int i = k + 1; i < N; ++i) {
            const double l = A[i][k] / A[k][k];
#pragma unroll
            for (int j = k; j < N; ++j) A[i][j] = fma(-l, A[k][j], A[i][j]);
            b[i] = fma(-l, b[k], b[i]);
        }
    }
    b[N - 1] = b[N - 1] / A[N - 1][N - 1];
#pragma unroll
    for (int k = N - 2; k >= 0; --k) {
#pragma unroll
        for (int j = k + 1; j < N; ++j) b[k] = fma(-A[k][j], b[j], b[k]);
        b[k] = b[k] / A[k][k];
    }
}

// compute_voronoi_vertex (geometry/core/voronoi.h:71-94): the centre of the
// sphere through the simplex's vertices, from 2 (p_i - p_0) . x = p_i . p_i -
// p_0 . p_0, i = 1..N.  One thread per simplex.
template <int N>
__global__ void __launch_bounds__(CC_BLOCK) k_voronoi(const double* __restrict__ pts, const int32_t* __restrict__ sim,
                                                       int64_t S, double* __restrict__ vv, int* __restrict__ bad)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const int32_t* sv = sim + s * (N + 1);
    double p0[N];
#pragma unroll
    for (int i = 0; i < N; ++i) p0[i] = pts[(int64_t)sv[0] * N + i];
    const double dot0 = dotn<N>(p0, p0);
    double A[N][N], b[N];
#pragma unroll
    for (int row = 0; row < N; ++row) {
        double p[N];
#pragma unroll
        for (int i = 0; i < N; ++i) p[i] = pts[(int64_t)sv[row + 1] * N + i];
#pragma unroll
        for (int col = 0; col < N; ++col) A[row][col] = 2 * (p[col] - p0[col]);
        b[row] = dotn<N>(p, p) - dot0;
    }
    gauss_solve<N>(A, b);
#pragma unroll
    for (int i = 0; i < N; ++i) vv[s * N + i] = b[i];
    if (!finiten<N>(b)) *bad = 1;
}

// ---------------------------------------------------------------- per-point data

// voronoi_edge_intersects_cocone (cocone.h:36-58)
__device__ inline bool edge_meets_cocone(double cos_a, double cos_b)
{
    if (fabs(cos_a) < COS_COCONE || fabs(cos_b) < COS_COCONE) return true;
    return (cos_a < 0 && cos_b > 0) || (cos_a > 0 && cos_b < 0);
}

__device__ inline bool inside_or_equal(double c) { return fabs(c) <= COS_COCONE; }

// intersect_cocone (cocone.h:84-157): the larger distance from the point of
// the points a + t ab, t >= 0, on the double cone of half-angle 3 pi / 8 about
// the axis, with the roots of quadratic_equation (numerical/quadratic.h:25-48).
template <int N>
__device__ inline bool cocone_distance(const double* axis, const double* a, const double* ab, double* distance)
{
    const double n_ab = dotn<N>(axis, ab);
    const double a_n = dotn<N>(a, axis);
    const double sq_a = dotn<N>(a, a);
    const double sq_ab = dotn<N>(ab, ab);
    const double a_ab = dotn<N>(a, ab);
    const double sq_cos = COS_COCONE * COS_COCONE;
    const double qa = n_ab * n_ab - sq_cos * sq_ab;
    const double qb = 2 * (a_n * n_ab - a_ab * sq_cos);
    const double qc = a_n * a_n - sq_a * sq_cos;
    const double disc = qb * qb - 4 * qa * qc;
    if (disc < 0) return false;
    const double sd = sqrt(disc);
    double t1, t2;
    if (qb >= 0) {
        t1 = (-qb - sd) / (2 * qa);
        t2 = 2 * qc / (-qb - sd);
    } else {
        t1 = (-qb + sd) / (2 * qa);
        t2 = 2 * qc / (-qb + sd);
    }
    const bool ok1 = t1 >= 0 && t1 <= DBL_MAX;
    const bool ok2 = t2 >= 0 && t2 <= DBL_MAX;
    if (!ok1 && !ok2) return false;
    double p1[N], p2[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        p1[i] = a[i] + t1 * ab[i];
        p2[i] = a[i] + t2 * ab[i];
    }
    if (ok1 && !ok2) {
        *distance = sqrt(dotn<N>(p1, p1));
        return true;
    }
    if (!ok1 && ok2) {
        *distance = sqrt(dotn<N>(p2, p2));
        return true;
    }
    const double d1 = dotn<N>(p1, p1), d2 = dotn<N>(p2, p2);
    *distance = d1 > d2 ? sqrt(d1) : sqrt(d2);
    return true;
}

struct VertexArgs {
    int64_t count;
    const double* pts;       // [count][N]
    const double* vv;        // [S][N] Voronoi vertices
    const int32_t* fs0;      // [F]
    const int32_t* fs1;      // [F], -1: one-sided
    const double* fortho;    // [F][N]
    const int64_t* vf_start; // [count + 1] incident facets per point, in facet order
    const int32_t* vf_fac;
    const int8_t* vf_loc;    // the point's position in the facet
    const int64_t* vo_start; // [count + 1] incident simplices per point, in simplex order
    const int32_t* vo_obj;
    double* pn;              // [count][N] positive pole direction
    double* height;
    double* radius;
    uint8_t* fcoc;           // [F][N] the facet's Voronoi edge meets the cocone of its vertex k
    int8_t* err;             // [count]
};

// vertex_and_facet_data (structure.cpp:343-398) for one point: the positive
// pole (voronoi_positive_norm, :61-104), with FIND_ALL the height
// (voronoi_height, :106-137) and radius of the Voronoi cell, and the cocone
// flags of its facets (cocone_facets_and_voronoi_radius, :139-268).
template <int N, bool FIND_ALL>
__global__ void __launch_bounds__(CC_BLOCK) k_vertex_data(VertexArgs a)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= a.count) return;
    double p[N];
#pragma unroll
    for (int i = 0; i < N; ++i) p[i] = a.pts[v * N + i];
    const int64_t f0 = a.vf_start[v], f1 = a.vf_start[v + 1];
    const int64_t o0 = a.vo_start[v], o1 = a.vo_start[v + 1];
    double pn[N];
#pragma unroll
    for (int i = 0; i < N; ++i) pn[i] = 0;
    auto finish = [&](int8_t e, double h, double r) {
#pragma unroll
        for (int i = 0; i < N; ++i) a.pn[v * N + i] = pn[i];
        a.height[v] = h;
        a.radius[v] = r;
        a.err[v] = e;
    };
    if (f0 == f1 && o0 == o1) {  // a dropped duplicate: in no simplex
        finish(VE_NONE, 0, 0);
        return;
    }
    bool unbounded = false;
    for (int64_t k = f0; k < f1; ++k) unbounded = unbounded || a.fs1[a.vf_fac[k]] < 0;
    if (unbounded) {
        for (int64_t k = f0; k < f1; ++k) {
            const int32_t f = a.vf_fac[k];
            if (a.fs1[f] >= 0) continue;
#pragma unroll
            for (int i = 0; i < N; ++i) pn[i] = pn[i] + a.fortho[(int64_t)f * N + i];
        }
    } else {
        double maxd = -DBL_MAX;
        for (int64_t k = o0; k < o1; ++k) {
            double vp[N];
#pragma unroll
            for (int i = 0; i < N; ++i) vp[i] = a.vv[(int64_t)a.vo_obj[k] * N + i] - p[i];
            const double d = dotn<N>(vp, vp);
            if (d > maxd) {
                maxd = d;
#pragma unroll
                for (int i = 0; i < N; ++i) pn[i] = vp[i];
            }
        }
    }
    normalizen<N>(pn);
    if (!finiten<N>(pn)) {
        finish(VE_POSITIVE_NOT_FINITE, 0, 0);
        return;
    }
    double height = 0;
    if (FIND_ALL) {
        double maxd = -DBL_MAX;
        bool found = false;
        for (int64_t k = o0; k < o1; ++k) {
            double vp[N];
#pragma unroll
            for (int i = 0; i < N; ++i) vp[i] = a.vv[(int64_t)a.vo_obj[k] * N + i] - p[i];
            if (dotn<N>(vp, pn) >= 0) continue;
            const double d = dotn<N>(vp, vp);
            if (d > maxd) {
                maxd = d;
                found = true;
            }
        }
        if (!found) {
            finish(VE_NEGATIVE_NOT_FOUND, 0, 0);
            return;
        }
        height = sqrt(maxd);
        if (!isfinite(height)) {
            finish(VE_NEGATIVE_NOT_FINITE, height, 0);
            return;
        }
    }
    double radius = 0;
    for (int64_t k = f0; k < f1; ++k) {
        const int32_t f = a.vf_fac[k];
        const int32_t s0 = a.fs0[f], s1 = a.fs1[f];
        const bool one = s1 < 0;
        double pa[N];
#pragma unroll
        for (int i = 0; i < N; ++i) pa[i] = a.vv[(int64_t)s0 * N + i] - p[i];
        const double pa_len = sqrt(dotn<N>(pa, pa));
        const double cos_a = dotn<N>(pn, pa) / pa_len;
        double pb_len, cos_b;
        if (one) {
            pb_len = 0;
            cos_b = dotn<N>(pn, a.fortho + (int64_t)f * N);
        } else {
            double pb[N];
#pragma unroll
            for (int i = 0; i < N; ++i) pb[i] = a.vv[(int64_t)s1 * N + i] - p[i];
            pb_len = sqrt(dotn<N>(pb, pb));
            cos_b = dotn<N>(pn, pb) / pb_len;
        }
        if (!edge_meets_cocone(cos_a, cos_b)) continue;
        a.fcoc[(int64_t)f * N + a.vf_loc[k]] = 1;
        if (!FIND_ALL || radius == DBL_MAX) continue;
        // voronoi_edge_radius (structure.cpp:139-213)
        double r;
        if (one && inside_or_equal(cos_b)) {
            r = DBL_MAX;
        } else if (!one && inside_or_equal(cos_a) && inside_or_equal(cos_b)) {
            r = pa_len > pb_len ? pa_len : pb_len;
        } else {
            double ab[N];
#pragma unroll
            for (int i = 0; i < N; ++i)
                ab[i] = one ? a.fortho[(int64_t)f * N + i] : a.vv[(int64_t)s1 * N + i] - a.vv[(int64_t)s0 * N + i];
            double maxd;
            if (!cocone_distance<N>(pn, pa, ab, &maxd)) {
                if (fabs(cos_a) > LIMIT_COS_PA_POLE) {
                    const double ab_len = one ? 1.0 : sqrt(dotn<N>(ab, ab));
                    const double cos_pa_ab = dotn<N>(pa, ab) / (pa_len * ab_len);
                    if (cos_pa_ab < LIMIT_COS_PA_AB) {
                        maxd = 0;
                    } else {
                        finish(VE_PA_CLOSE, height, radius);
                        return;
                    }
                } else {
                    finish(VE_PA_FAR, height, radius);
                    return;
                }
            }
            if (!isfinite(maxd)) {
                finish(VE_DISTANCE_NOT_FINITE, height, radius);
                return;
            }
            r = inside_or_equal(cos_a) ? (pa_len > maxd ? pa_len : maxd) : maxd;
        }
        if (r > radius) radius = r;
    }
    if (FIND_ALL && !(radius > 0 && radius <= DBL_MAX)) {
        finish(VE_RADIUS, height, radius);
        return;
    }
    finish(VE_NONE, height, radius);
}

// ---------------------------------------------------------------- BoundCocone interior points

struct InteriorArgs {
    int64_t count;
    const double* pn;
    const double* height;
    const double* radius;
    const int64_t* vf_start;
    const int32_t* vf_fac;
    const int8_t* vf_loc;
    const int32_t* fv;
    const uint8_t* fcoc;
    double rho, cos_alpha;
    uint8_t* interior;
    int* flag;
};

// find_interior_vertices (reconstruction.cpp:72-160), first rule: a thin
// Voronoi cell (radius <= rho * height) whose pole agrees with the poles of
// all its cocone neighbours (|cos| >= cos(alpha)).  The cocone neighbours of
// a point are the other vertices k of its facets with the flag of (facet, k)
// set (cocone_neighbors, structure.cpp:270-302); repeats do not change the rule.
template <int N>
__global__ void __launch_bounds__(CC_BLOCK) k_interior_first(InteriorArgs a)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= a.count) return;
    uint8_t in = 0;
    if (a.radius[v] <= a.rho * a.height[v]) {
        double pv[N];
#pragma unroll
        for (int i = 0; i < N; ++i) pv[i] = a.pn[v * N + i];
        bool flat = true;
        for (int64_t k = a.vf_start[v]; k < a.vf_start[v + 1] && flat; ++k) {
            const int64_t f = a.vf_fac[k];
            const int loc = a.vf_loc[k];
#pragma unroll
            for (int u = 0; u < N; ++u) {
                if (u == loc || !a.fcoc[f * N + u]) continue;
                const int64_t w = a.fv[f * N + u];
                if (!(fabs(dotn<N>(pv, a.pn + w * N)) >= a.cos_alpha)) flat = false;
            }
        }
        in = flat ? 1 : 0;
    }
    a.interior[v] = in;
    if (in) *a.flag = 1;
}

// The second rule, applied until nothing changes: a thin cell whose pole
// agrees with the pole of an interior cocone neighbour.  The set grows
// monotonically to the same fixed point in any order of updates, so a point
// may read a neighbour's flag before or after this launch sets it.
template <int N>
__global__ void __launch_bounds__(CC_BLOCK) k_interior_grow(InteriorArgs a)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= a.count) return;
    if (a.interior[v] || !(a.radius[v] <= a.rho * a.height[v])) return;
    double pv[N];
#pragma unroll
    for (int i = 0; i < N; ++i) pv[i] = a.pn[v * N + i];
    for (int64_t k = a.vf_start[v]; k < a.vf_start[v + 1]; ++k) {
        const int64_t f = a.vf_fac[k];
        const int loc = a.vf_loc[k];
#pragma unroll
        for (int u = 0; u < N; ++u) {
            if (u == loc || !a.fcoc[f * N + u]) continue;
            const int64_t w = a.fv[f * N + u];
            if (!((volatile const uint8_t*)a.interior)[w]) continue;
            if (fabs(dotn<N>(pv, a.pn + w * N)) >= a.cos_alpha) {
                a.interior[v] = 1;
                *a.flag = 1;
                return;
            }
        }
    }
}

// Facets to start from: Cocone keeps a facet whose Voronoi edge meets the
// cocones of all its vertices (find_cocone_facets, reconstruction.cpp:49-70);
// BoundCocone one whose interior vertices all have the flag, with at least one
// interior vertex (cocone_interior_facet, :162-181).
template <int N, bool BOUND>
__global__ void __launch_bounds__(CC_BLOCK) k_keep(int64_t F, const uint8_t* __restrict__ fcoc,
                                                    const int32_t* __restrict__ fv, const uint8_t* __restrict__ interior,
                                                    uint8_t* __restrict__ keep, int* __restrict__ flag)
{
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    bool k;
    if (!BOUND) {
        k = true;
#pragma unroll
        for (int u = 0; u < N; ++u) k = k && fcoc[f * N + u];
    } else {
        bool ok = true, found = false;
#pragma unroll
        for (int u = 0; u < N; ++u) {
            const bool in = interior[fv[f * N + u]];
            const bool ic = in && fcoc[f * N + u];
            ok = ok && (ic || !in);
            found = found || ic;
        }
        k = ok && found;
    }
    keep[f] = k ? 1 : 0;
    if (k) *flag = 1;
}

// ---------------------------------------------------------------- host

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

unsigned blocks_for(int64_t n) { return (unsigned)((n + CC_BLOCK - 1) / CC_BLOCK); }

template <typename T>
void upload(DevBuf& d, const std::vector<T>& h, hipStream_t s)
{
    d.alloc(sizeof(T) * std::max<size_t>(h.size(), 1));
    if (!h.empty()) NSP_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
}

std::string not_reconstructable(const char* what, int n)
{
    return std::string(what) + ". " + std::to_string(n - 1) + "-manifold is not reconstructable.";
}

// Facets: every face of every simplex, grouped by sorted vertex tuple
// (create_delaunay_objects_and_facets, geometry/core/delaunay.h:86-140).  The
// first simplex of a facet is the one with the smaller index and the facet's
// normal is that simplex's normal of the face, pointing away from it.
template <int N>
void build_facets(nsp_manifold* m, const std::vector<double>& orthos, std::vector<double>* fortho)
{
    struct Face {
        std::array<int32_t, N> v;
        int32_t s;
        int8_t r;
    };
    std::vector<Face> faces((size_t)m->S * (N + 1));
    for (int64_t s = 0; s < m->S; ++s) {
        const int32_t* sv = &m->simplices[(size_t)s * (N + 1)];
        for (int r = 0; r <= N; ++r) {
            Face& x = faces[(size_t)s * (N + 1) + r];
            int c = 0;
            for (int i = 0; i <= N; ++i)
                if (i != r) x.v[c++] = sv[i];
            std::sort(x.v.begin(), x.v.end());
            x.s = (int32_t)s;
            x.r = (int8_t)r;
        }
    }
    std::sort(faces.begin(), faces.end(), [](const Face& a, const Face& b) {
        if (a.v != b.v) return a.v < b.v;
        return a.s < b.s;
    });
    m->fv.clear();
    m->fs0.clear();
    m->fs1.clear();
    fortho->clear();
    for (size_t i = 0; i < faces.size();) {
        size_t j = i + 1;
        while (j < faces.size() && faces[j].v == faces[i].v) ++j;
        if (j - i > 2) raise(NSP_E_INVALID, "a Delaunay facet in more than two simplices");
        m->fv.insert(m->fv.end(), faces[i].v.begin(), faces[i].v.end());
        m->fs0.push_back(faces[i].s);
        m->fs1.push_back(j - i == 2 ? faces[i + 1].s : -1);
        const double* o = &orthos[((size_t)faces[i].s * (N + 1) + faces[i].r) * N];
        fortho->insert(fortho->end(), o, o + N);
        i = j;
    }
    m->F = (int64_t)m->fs0.size();
}

// CSR lists of (facet, position) per point in facet order and of simplices
// per point in simplex order (vertex_connections, structure.cpp:304-325).
template <int N>
void build_connections(const nsp_manifold* m, std::vector<int64_t>* vf_start, std::vector<int32_t>* vf_fac,
                       std::vector<int8_t>* vf_loc, std::vector<int64_t>* vo_start, std::vector<int32_t>* vo_obj)
{
    const int64_t P = m->count;
    vf_start->assign(P + 1, 0);
    vo_start->assign(P + 1, 0);
    for (int32_t v : m->fv) (*vf_start)[v + 1]++;
    for (int32_t v : m->simplices) (*vo_start)[v + 1]++;
    for (int64_t i = 0; i < P; ++i) {
        (*vf_start)[i + 1] += (*vf_start)[i];
        (*vo_start)[i + 1] += (*vo_start)[i];
    }
    vf_fac->resize((size_t)(*vf_start)[P]);
    vf_loc->resize((size_t)(*vf_start)[P]);
    vo_obj->resize((size_t)(*vo_start)[P]);
    std::vector<int64_t> fill(vf_start->begin(), vf_start->end() - 1);
    for (int64_t f = 0; f < m->F; ++f)
        for (int k = 0; k < N; ++k) {
            const int64_t at = fill[m->fv[(size_t)f * N + k]]++;
            (*vf_fac)[at] = (int32_t)f;
            (*vf_loc)[at] = (int8_t)k;
        }
    fill.assign(vo_start->begin(), vo_start->end() - 1);
    for (int64_t s = 0; s < m->S; ++s)
        for (int k = 0; k <= N; ++k) (*vo_obj)[fill[m->simplices[(size_t)s * (N + 1) + k]]++] = (int32_t)s;
}

template <int N>
void create_t(nsp_manifold* m, const std::vector<double>& orthos)
{
    const hipStream_t st = m->stream;
    double t0 = now_ms();
    std::vector<double> fortho;
    build_facets<N>(m, orthos, &fortho);
    std::vector<int64_t> vf_start, vo_start;
    std::vector<int32_t> vf_fac, vo_obj;
    std::vector<int8_t> vf_loc;
    build_connections<N>(m, &vf_start, &vf_fac, &vf_loc, &vo_start, &vo_obj);
    double t1 = now_ms();
    m->ms[1] = t1 - t0;

    DevBuf d_pts, d_sim, d_vv, d_fs0, d_fs1, d_fortho, d_vo_start, d_vo_obj, d_err;
    upload(d_pts, m->points, st);
    upload(d_sim, m->simplices, st);
    upload(m->d_fv, m->fv, st);
    upload(d_fs0, m->fs0, st);
    upload(d_fs1, m->fs1, st);
    upload(d_fortho, fortho, st);
    upload(m->d_vf_start, vf_start, st);
    upload(m->d_vf_fac, vf_fac, st);
    upload(m->d_vf_loc, vf_loc, st);
    upload(d_vo_start, vo_start, st);
    upload(d_vo_obj, vo_obj, st);
    const int64_t P = m->count;
    d_vv.alloc(sizeof(double) * std::max<int64_t>(m->S * N, 1));
    m->d_fcoc.alloc(std::max<int64_t>(m->F * N, 1));
    m->d_pn.alloc(sizeof(double) * P * N);
    m->d_height.alloc(sizeof(double) * P);
    m->d_radius.alloc(sizeof(double) * P);
    m->d_interior.alloc(P);
    m->d_keep.alloc(std::max<int64_t>(m->F, 1));
    m->d_flag.alloc(sizeof(int));
    d_err.alloc(P);
    NSP_HIP(hipMemsetAsync(m->d_flag.p, 0, sizeof(int), st));
    NSP_HIP(hipMemsetAsync(m->d_fcoc.p, 0, std::max<int64_t>(m->F * N, 1), st));
    if (m->S > 0)
        k_voronoi<N><<<blocks_for(m->S), CC_BLOCK, 0, st>>>(d_pts.as<double>(), d_sim.as<int32_t>(), m->S,
                                                            d_vv.as<double>(), m->d_flag.as<int>());
    NSP_HIP(hipGetLastError());
    VertexArgs va;
    va.count = P;
    va.pts = d_pts.as<double>();
    va.vv = d_vv.as<double>();
    va.fs0 = d_fs0.as<int32_t>();
    va.fs1 = d_fs1.as<int32_t>();
    va.fortho = d_fortho.as<double>();
    va.vf_start = m->d_vf_start.as<int64_t>();
    va.vf_fac = m->d_vf_fac.as<int32_t>();
    va.vf_loc = m->d_vf_loc.as<int8_t>();
    va.vo_start = d_vo_start.as<int64_t>();
    va.vo_obj = d_vo_obj.as<int32_t>();
    va.pn = m->d_pn.as<double>();
    va.height = m->d_height.as<double>();
    va.radius = m->d_radius.as<double>();
    va.fcoc = m->d_fcoc.as<uint8_t>();
    va.err = d_err.as<int8_t>();
    if (m->cocone_only)
        k_vertex_data<N, false><<<blocks_for(P), CC_BLOCK, 0, st>>>(va);
    else
        k_vertex_data<N, true><<<blocks_for(P), CC_BLOCK, 0, st>>>(va);
    NSP_HIP(hipGetLastError());
    int bad = 0;
    std::vector<int8_t> err(P);
    m->pn.resize((size_t)P * N);
    NSP_HIP(hipMemcpyAsync(&bad, m->d_flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
    NSP_HIP(hipMemcpyAsync(err.data(), d_err.p, P, hipMemcpyDeviceToHost, st));
    NSP_HIP(hipMemcpyAsync(m->pn.data(), m->d_pn.p, sizeof(double) * P * N, hipMemcpyDeviceToHost, st));
    NSP_HIP(hipStreamSynchronize(st));
    m->ms[2] = now_ms() - t1;
    if (bad) raise(NSP_E_INVALID, "Voronoi vertex is not finite");
    for (int64_t v = 0; v < P; ++v)
        if (err[v]) raise(NSP_E_INVALID, vertex_error_text(err[v]));
}

// ---------------------------------------------------------------- pruning and extraction

// sharp_ridge (prune_facets.h:56-115): the ridge's live facets e (insertion
// order) lie in a wedge of less than pi / 2 about it; a ridge with a
// non-interior vertex is never sharp, a ridge of one facet always is.
template <int N>
bool sharp_ridge(const nsp_manifold* m, const uint8_t* interior, const int32_t* rv, const int32_t* points, int cnt)
{
    for (int i = 0; i < N - 1; ++i)
        if (!interior[rv[i]]) return false;
    if (cnt == 1) return true;
    const double* P = m->points.data();
    auto diff = [&](int32_t a, double* out) {
        for (int c = 0; c < N; ++c) out[c] = P[(int64_t)a * N + c] - P[(int64_t)rv[0] * N + c];
    };
    double vecs[N * N], e0[N], e1[N];
    for (int i = 0; i < N - 2; ++i) diff(rv[i + 1], vecs + i * N);
    diff(points[0], vecs + (N - 2) * N);
    ortho_nn<N, double>(vecs, e0);
    normalizen<N>(e0);
    std::memcpy(vecs + (N - 2) * N, e0, sizeof(double) * N);
    ortho_nn<N, double>(vecs, e1);
    normalizen<N>(e1);
    double bv[N];
    diff(points[0], bv);
    double base[2] = {dotn<N>(e0, bv), dotn<N>(e1, bv)};
    normalizen<2>(base);
    double cos_plus = 1, cos_minus = 1, sin_plus = 0, sin_minus = 0;
    for (int k = 1; k < cnt; ++k) {
        double fv[N];
        diff(points[k], fv);
        double w[2] = {dotn<N>(e0, fv), dotn<N>(e1, fv)};
        normalizen<2>(w);
        const double sine = base[0] * w[1] - base[1] * w[0];
        const double cosine = dotn<2>(base, w);
        if (sine >= 0) {
            if (cosine < cos_plus) {
                cos_plus = cosine;
                sin_plus = sine;
            }
        } else {
            if (cosine < cos_minus) {
                cos_minus = cosine;
                sin_minus = sine;
            }
        }
    }
    if (cos_plus <= 0 || cos_minus <= 0) return false;
    return cos_plus * cos_minus - fabs(sin_plus * sin_minus) > 0;
}

// prune_facets_incident_to_sharp_ridges (prune_facets.h:117-200): rounds over
// suspicious ridges (first every ridge of a kept facet, then the other ridges
// of the facets removed in the round before) in sorted order; a sharp ridge
// removes all its facets.
template <int N>
void prune(const nsp_manifold* m, const uint8_t* interior, std::vector<uint8_t>& keep)
{
    constexpr int K = N - 1;
    using RidgeV = std::array<int32_t, K>;
    struct Rec {
        RidgeV v;
        int32_t facet;
        int32_t point;  // the facet's vertex outside the ridge
        int64_t order;
    };
    auto ridge_of = [&](int64_t f, int r) {
        RidgeV x;
        int c = 0;
        for (int q = 0; q < N; ++q)
            if (q != r) x[c++] = m->fv[(size_t)f * N + q];
        return x;
    };
    std::vector<Rec> R;
    for (int64_t f = 0; f < m->F; ++f) {
        if (!keep[f]) continue;
        for (int r = 0; r < N; ++r) {
            const int64_t order = (int64_t)R.size();
            R.push_back({ridge_of(f, r), (int32_t)f, m->fv[(size_t)f * N + r], order});
        }
    }
    std::sort(R.begin(), R.end(), [](const Rec& a, const Rec& b) {
        if (a.v != b.v) return a.v < b.v;
        return a.order < b.order;
    });
    std::vector<uint8_t> alive(R.size(), 1);
    auto find = [&](const RidgeV& v) {
        return (size_t)(std::lower_bound(R.begin(), R.end(), v, [](const Rec& a, const RidgeV& b) { return a.v < b; }) -
                        R.begin());
    };
    std::vector<RidgeV> sus, next;
    for (size_t i = 0; i < R.size(); ++i)
        if (i == 0 || R[i].v != R[i - 1].v) sus.push_back(R[i].v);
    std::vector<size_t> e;
    std::vector<int32_t> pts;
    while (!sus.empty()) {
        next.clear();
        for (const RidgeV& rv : sus) {
            e.clear();
            pts.clear();
            for (size_t i = find(rv); i < R.size() && R[i].v == rv; ++i)
                if (alive[i]) {
                    e.push_back(i);
                    pts.push_back(R[i].point);
                }
            if (e.empty()) continue;
            if (!sharp_ridge<N>(m, interior, rv.data(), pts.data(), (int)e.size())) continue;
            for (size_t i : e) {
                const int64_t f = R[i].facet;
                for (int r = 0; r < N; ++r)
                    if (m->fv[(size_t)f * N + r] != R[i].point) next.push_back(ridge_of(f, r));
                keep[f] = 0;
            }
            for (size_t i : e) {
                const int64_t f = R[i].facet;
                for (int r = 0; r < N; ++r) {
                    const RidgeV x = ridge_of(f, r);
                    for (size_t j = find(x); j < R.size() && R[j].v == x; ++j)
                        if (R[j].facet == f) alive[j] = 0;
                }
            }
        }
        std::sort(next.begin(), next.end());
        next.erase(std::unique(next.begin(), next.end()), next.end());
        sus.swap(next);
    }
}

// extract_manifold (extract_manifold.h:50-131): walk the simplices reachable
// from the outside (the one-sided facets) without crossing a kept facet; the
// kept facets met on the way are the manifold.
void extract(const nsp_manifold* m, std::vector<uint8_t>& keep)
{
    const int64_t S = m->S, F = m->F;
    std::vector<int64_t> start(S + 1, 0);
    for (int64_t f = 0; f < F; ++f) {
        start[m->fs0[f] + 1]++;
        if (m->fs1[f] >= 0) start[m->fs1[f] + 1]++;
    }
    for (int64_t s = 0; s < S; ++s) start[s + 1] += start[s];
    std::vector<int32_t> of((size_t)start[S]);
    std::vector<int64_t> fill(start.begin(), start.end() - 1);
    for (int64_t f = 0; f < F; ++f) {
        of[fill[m->fs0[f]]++] = (int32_t)f;
        if (m->fs1[f] >= 0) of[fill[m->fs1[f]]++] = (int32_t)f;
    }
    std::vector<uint8_t> vis(S, 0), met(F, 0);
    std::vector<int32_t> stack;
    for (int64_t f = 0; f < F; ++f)
        if (m->fs1[f] < 0) stack.push_back((int32_t)f);
    while (!stack.empty()) {
        const int32_t f = stack.back();
        stack.pop_back();
        if (keep[f]) {
            met[f] = 1;
            continue;
        }
        const int32_t s0 = m->fs0[f], s1 = m->fs1[f];
        int32_t d;
        if (s1 < 0) {
            if (vis[s0]) continue;
            d = s0;
        } else {
            if (vis[s0] && vis[s1]) continue;
            d = vis[s0] ? s1 : s0;
        }
        vis[d] = 1;
        for (int64_t k = start[d]; k < start[d + 1]; ++k)
            if (of[k] != f) stack.push_back(of[k]);
    }
    keep.swap(met);
}

// The common part of cocone() and bound_cocone() (reconstruction.cpp:285-319).
template <int N>
int64_t reconstruct_t(nsp_manifold* m, bool bound, double rho, double alpha)
{
    const hipStream_t st = m->stream;
    const double t0 = now_ms();
    const int64_t P = m->count, F = m->F;
    std::vector<uint8_t> interior(P, 1);
    NSP_HIP(hipMemsetAsync(m->d_flag.p, 0, sizeof(int), st));
    int flag = 0;
    if (bound) {
        InteriorArgs ia;
        ia.count = P;
        ia.pn = m->d_pn.as<double>();
        ia.height = m->d_height.as<double>();
        ia.radius = m->d_radius.as<double>();
        ia.vf_start = m->d_vf_start.as<int64_t>();
        ia.vf_fac = m->d_vf_fac.as<int32_t>();
        ia.vf_loc = m->d_vf_loc.as<int8_t>();
        ia.fv = m->d_fv.as<int32_t>();
        ia.fcoc = m->d_fcoc.as<uint8_t>();
        ia.rho = rho;
        ia.cos_alpha = std::cos(alpha);
        ia.interior = m->d_interior.as<uint8_t>();
        ia.flag = m->d_flag.as<int>();
        k_interior_first<N><<<blocks_for(P), CC_BLOCK, 0, st>>>(ia);
        NSP_HIP(hipGetLastError());
        NSP_HIP(hipMemcpyAsync(&flag, m->d_flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
        NSP_HIP(hipStreamSynchronize(st));
        if (!flag) raise(NSP_E_INVALID, not_reconstructable("Interior vertices not found", N));
        // grow until a launch sets nothing; each launch adds at least one point
        for (int64_t round = 0; round <= P; ++round) {
            NSP_HIP(hipMemsetAsync(m->d_flag.p, 0, sizeof(int), st));
            k_interior_grow<N><<<blocks_for(P), CC_BLOCK, 0, st>>>(ia);
            NSP_HIP(hipGetLastError());
            NSP_HIP(hipMemcpyAsync(&flag, m->d_flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
            NSP_HIP(hipStreamSynchronize(st));
            if (!flag) break;
        }
        NSP_HIP(hipMemsetAsync(m->d_flag.p, 0, sizeof(int), st));
        k_keep<N, true><<<blocks_for(F), CC_BLOCK, 0, st>>>(F, m->d_fcoc.as<uint8_t>(), m->d_fv.as<int32_t>(),
                                                            m->d_interior.as<uint8_t>(), m->d_keep.as<uint8_t>(),
                                                            m->d_flag.as<int>());
        NSP_HIP(hipMemcpyAsync(interior.data(), m->d_interior.p, P, hipMemcpyDeviceToHost, st));
    } else {
        k_keep<N, false><<<blocks_for(F), CC_BLOCK, 0, st>>>(F, m->d_fcoc.as<uint8_t>(), m->d_fv.as<int32_t>(),
                                                             nullptr, m->d_keep.as<uint8_t>(), m->d_flag.as<int>());
    }
    NSP_HIP(hipGetLastError());
    std::vector<uint8_t> keep(F);
    NSP_HIP(hipMemcpyAsync(&flag, m->d_flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
    NSP_HIP(hipMemcpyAsync(keep.data(), m->d_keep.p, F, hipMemcpyDeviceToHost, st));
    NSP_HIP(hipStreamSynchronize(st));
    const double t1 = now_ms();
    m->ms[3] = t1 - t0;
    if (!flag)
        raise(NSP_E_INVALID,
              not_reconstructable(bound ? "Cocone interior facets not found" : "Cocone facets not found", N));
    prune<N>(m, interior.data(), keep);
    const double t2 = now_ms();
    m->ms[4] = t2 - t1;
    if (std::find(keep.begin(), keep.end(), 1) == keep.end())
        raise(NSP_E_INVALID, not_reconstructable("Cocone facets not found after prune", N));
    extract(m, keep);
    int64_t k = 0;
    for (int64_t f = 0; f < F; ++f) k += keep[f];
    if (k == 0) raise(NSP_E_INVALID, not_reconstructable("Cocone facets not found after manifold extraction", N));
    // create_normals_and_facets (reconstruction.cpp:220-250)
    m->out_facets.clear();
    m->out_facets.reserve((size_t)k * N);
    m->out_normals.assign((size_t)P * N, 0.0);
    for (int64_t f = 0; f < F; ++f) {
        if (!keep[f]) continue;
        for (int j = 0; j < N; ++j) {
            const int32_t p = m->fv[(size_t)f * N + j];
            m->out_facets.push_back(p);
            std::memcpy(&m->out_normals[(size_t)p * N], &m->pn[(size_t)p * N], sizeof(double) * N);
        }
    }
    m->ms[5] = now_ms() - t2;
    return k;
}

template <typename Fn>
int guarded(std::string* err, Fn&& fn)
{
    try {
        fn();
        return NSP_OK;
    } catch (const Error& e) {
        if (err) *err = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        if (err) *err = "host allocation failed";
        return NSP_E_NOMEM;
    }
}

}  // namespace
}  // namespace nsp

using namespace nsp;

int nsp_manifold_create(nsp_context* ctx, int n, const float* points, int64_t count, int cocone_only,
                        nsp_manifold** out)
{
    if (!ctx || !out) return NSP_E_INVALID;
    *out = nullptr;
    std::string* err = context_error(ctx);
    return guarded(err, [&] {
        if (n < 2 || n > 5) raise(NSP_E_UNSUPPORTED, "manifold reconstruction dimension must be in [2, 5]");
        if (!points && count > 0) raise(NSP_E_INVALID, "points is NULL");
        // ManifoldConstructorImpl (reconstruction.cpp:392-406)
        if (count < n + 2)
            raise(NSP_E_INVALID, "Error point count " + std::to_string(count) +
                                     " for cocone manifold reconstruction in " + std::to_string(n) + "-space");
        if (count > INT32_MAX) raise(NSP_E_INVALID, "too many points");
        auto m = std::make_unique<nsp_manifold>();
        m->n = n;
        m->count = count;
        m->cocone_only = cocone_only != 0;
        m->device = context_device(ctx);
        m->stream = context_stream(ctx);
        const double t0 = now_ms();
        nsp_hull* h = nullptr;
        const int rc = nsp_delaunay(ctx, n, points, count, 0, &h);
        if (rc != NSP_OK) raise(rc, *err);
        std::unique_ptr<nsp_hull, void (*)(nsp_hull*)> hold(h, nsp_hull_destroy);
        int64_t S = 0;
        nsp_hull_info(h, nullptr, &S, nullptr, nullptr, nullptr);
        if (S == 0) raise(NSP_E_INVALID, "no Delaunay simplices");
        if (S * (n + 1) > INT32_MAX) raise(NSP_E_INVALID, "too many Delaunay simplices");
        m->S = S;
        m->simplices.resize((size_t)S * (n + 1));
        m->points.resize((size_t)count * n);
        std::vector<double> orthos((size_t)S * (n + 1) * n);
        nsp_hull_copy(h, m->simplices.data(), orthos.data(), m->points.data(), nullptr);
        m->ms[0] = now_ms() - t0;
        NSP_HIP(hipSetDevice(m->device));
        switch (n) {
        case 2: create_t<2>(m.get(), orthos); break;
        case 3: create_t<3>(m.get(), orthos); break;
        case 4: create_t<4>(m.get(), orthos); break;
        case 5: create_t<5>(m.get(), orthos); break;
        }
        *out = m.release();
    });
}

static int manifold_run(nsp_manifold* m, bool bound, double rho, double alpha, int64_t* facets, std::string* err)
{
    return guarded(err, [&] {
        NSP_HIP(hipSetDevice(m->device));
        if (bound) {
            if (m->cocone_only) raise(NSP_E_INVALID, "Manifold constructor created for Cocone and not for BoundCocone");
            // check_rho_and_aplha (reconstruction.cpp:265-278)
            char buf[64];
            if (!(rho > 0 && rho < 1)) {
                std::snprintf(buf, sizeof buf, "%.10g", rho);
                raise(NSP_E_INVALID, std::string("Rho must be in the interval (0, 1), but rho = ") + buf);
            }
            if (!(alpha > 0 && alpha < 1)) {
                std::snprintf(buf, sizeof buf, "%.10g", alpha);
                raise(NSP_E_INVALID, std::string("Alpha must be in the interval (0, 1), but alpha = ") + buf);
            }
        }
        m->out_facets.clear();
        m->out_normals.clear();
        int64_t k = 0;
        switch (m->n) {
        case 2: k = reconstruct_t<2>(m, bound, rho, alpha); break;
        case 3: k = reconstruct_t<3>(m, bound, rho, alpha); break;
        case 4: k = reconstruct_t<4>(m, bound, rho, alpha); break;
        case 5: k = reconstruct_t<5>(m, bound, rho, alpha); break;
        }
        if (facets) *facets = k;
    });
}

int nsp_manifold_cocone(nsp_context* ctx, nsp_manifold* m, int64_t* facets)
{
    if (!ctx || !m) return NSP_E_INVALID;
    return manifold_run(m, false, 0, 0, facets, context_error(ctx));
}

int nsp_manifold_bound_cocone(nsp_context* ctx, nsp_manifold* m, double rho, double alpha, int64_t* facets)
{
    if (!ctx || !m) return NSP_E_INVALID;
    return manifold_run(m, true, rho, alpha, facets, context_error(ctx));
}

int nsp_manifold_result(const nsp_manifold* m, double* normals, int32_t* facets)
{
    if (!m) return NSP_E_INVALID;
    if (normals) std::copy(m->out_normals.begin(), m->out_normals.end(), normals);
    if (facets) std::copy(m->out_facets.begin(), m->out_facets.end(), facets);
    return NSP_OK;
}

int nsp_manifold_info(const nsp_manifold* m, int* n, int64_t* points, int64_t* delaunay_objects,
                      int64_t* delaunay_facets)
{
    if (!m) return NSP_E_INVALID;
    if (n) *n = m->n;
    if (points) *points = m->count;
    if (delaunay_objects) *delaunay_objects = m->S;
    if (delaunay_facets) *delaunay_facets = m->F;
    return NSP_OK;
}

int nsp_manifold_objects(const nsp_manifold* m, int32_t* objects)
{
    if (!m) return NSP_E_INVALID;
    if (objects) std::copy(m->simplices.begin(), m->simplices.end(), objects);
    return NSP_OK;
}

int nsp_manifold_stats(const nsp_manifold* m, double* ms)
{
    if (!m) return NSP_E_INVALID;
    if (ms)
        for (int i = 0; i < 6; ++i) ms[i] = m->ms[i];
    return NSP_OK;
}

void nsp_manifold_destroy(nsp_manifold* m) { delete m; }
