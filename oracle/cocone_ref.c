/*
 * Test infrastructure: the FP64 CPU checker of the Cocone and BoundCocone
 * surface reconstructions (templeblock/math src/geometry/cocone/), restated
 * step by step.  Only the tests and the benchmark's CPU-baseline leg use it;
 * the library never does.
 *
 * Input: a Delaunay triangulation (compute_delaunay's output, pinned by its own
 * checker in hull_ref.py): the points as doubles, the simplices' vertex indices
 * and, per simplex and vertex r, the unit normal of the face without vertex r
 * pointing away from vertex r (DelaunaySimplex::ortho).
 *
 * The reference keeps its Delaunay facets and its sets of suspicious ridges in
 * hash containers, so its iteration orders are the standard library's.  Here
 * the facets are ordered by their sorted vertex tuples, a facet's first
 * simplex is the one with the smaller index, and the suspicious ridges of
 * each pruning round are taken in sorted order; the library does the same.
 * Compiled with -ffp-contract=off: only the explicit fma() of the reference's
 * dot() and Gaussian elimination is fused.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "nsp_oracle.h"

#define CM 6

/* cos(3 pi / 8), COS_OF_AN_OPENING_ANGLE_WITH_THE_AXIS (cocone.h:29-31) */
static const double COS_COCONE = 0.38268343236508977172845998403039886676134456248562704143380063562754603396;
/* structure.cpp:30-32 */
static const double LIMIT_COSINE_FOR_INTERSECTION_PA_POLE = 0.99;
static const double LIMIT_COSINE_FOR_INTERSECTION_PA_AB = -0.9999;

void orc_set_error(const char* msg);

static double dot_c(const double* a, const double* b, int n)
{
    double r = a[0] * b[0];
    for (int i = 1; i < n; ++i) r = fma(a[i], b[i], r);
    return r;
}

static double length_c(const double* a, int n) { return sqrt(dot_c(a, a, n)); }

static void normalize_c(double* v, int n)
{
    const double len = length_c(v, n);
    for (int i = 0; i < n; ++i) v[i] = v[i] / len;
}

static int finite_c(const double* v, int n)
{
    for (int i = 0; i < n; ++i)
        if (!isfinite(v[i])) return 0;
    return 1;
}

/* solve_gauss (numerical/gauss.h:36-100): partial pivoting by the largest
 * absolute value below the diagonal (the first one on ties), fma updates. */
static void solve_gauss(int n, double* A, double* b)
{
    for (int k = 0; k < n - 1; ++k) {
        double mx = fabs(A[k * n + k]);
        int pivot = k;
        for (int r = k + 1; r < n; ++r) {
            const double v = fabs(A[r * n + k]);
            if (v > mx) {
                mx = v;
                pivot = r;
            }
        }
        if (pivot != k) {
            for (int j = 0; j < n; ++j) {
                const double t = A[pivot * n + j];
                A[pivot * n + j] = A[k * n + j];
                A[k * n + j] = t;
            }
            const double t = b[pivot];
            b[pivot] = b[k];
            b[k] = t;
        }
        for (int i = k + 1; i < n; ++i) {
            const double l = A[i * n + k] / A[k * n + k];
            for (int j = k; j < n; ++j) A[i * n + j] = fma(-l, A[k * n + j], A[i * n + j]);
            b[i] = fma(-l, b[k], b[i]);
        }
    }
    b[n - 1] = b[n - 1] / A[(n - 1) * n + n - 1];
    for (int k = n - 2; k >= 0; --k) {
        for (int j = k + 1; j < n; ++j) b[k] = fma(-A[k * n + j], b[j], b[k]);
        b[k] = b[k] / A[k * n + k];
    }
}

/* compute_voronoi_vertex (geometry/core/voronoi.h:71-94): the centre of the
 * sphere through the simplex's n + 1 vertices, from the n equations
 * 2 (p_i - p_0) . x = p_i . p_i - p_0 . p_0. */
void orc_voronoi_vertex(int n, const double* points, const int32_t* simplex, double* out)
{
    double A[CM * CM], b[CM];
    const double* p0 = points + (int64_t)simplex[0] * n;
    const double dot0 = dot_c(p0, p0, n);
    for (int row = 0; row < n; ++row) {
        const double* pn = points + (int64_t)simplex[row + 1] * n;
        for (int col = 0; col < n; ++col) A[row * n + col] = 2 * (pn[col] - p0[col]);
        b[row] = dot_c(pn, pn, n) - dot0;
    }
    solve_gauss(n, A, b);
    memcpy(out, b, sizeof(double) * n);
}

/* ------------------------------------------------------------------------ */
/* Delaunay facets (create_delaunay_objects_and_facets, delaunay.h:77-140)     */

typedef struct {
    int32_t v[CM];   /* sorted vertex indices (n of them) */
    int64_t s0, s1;  /* incident simplices (s1 < 0: one-sided) */
    double ortho[CM];
    int cocone[CM];  /* cocone_vertex (structure.h ManifoldFacet) */
} cfacet;

typedef struct {
    int32_t v[CM];
    int64_t s;
    int r;
} face_rec;

static int g_n;

static int cmp_face(const void* a, const void* b)
{
    const face_rec* x = a;
    const face_rec* y = b;
    for (int i = 0; i < g_n; ++i) {
        if (x->v[i] != y->v[i]) return x->v[i] < y->v[i] ? -1 : 1;
    }
    if (x->s != y->s) return x->s < y->s ? -1 : 1;
    return 0;
}

static void sort_ints(int32_t* a, int k)
{
    for (int i = 1; i < k; ++i) {
        int32_t t = a[i];
        int j = i - 1;
        while (j >= 0 && a[j] > t) {
            a[j + 1] = a[j];
            --j;
        }
        a[j + 1] = t;
    }
}

/* per vertex: incident facets (facet, local index) and objects */
typedef struct {
    int64_t* facet;
    int* local;
    int64_t nf, cf;
    int64_t* obj;
    int64_t no, co;
    int32_t* nbr;  /* cocone neighbours */
    int64_t nn, cn;
    double pn[CM];
    double height, radius;
} cvertex;

static void push_i64(int64_t** a, int64_t* n, int64_t* cap, int64_t v)
{
    if (*n == *cap) {
        *cap = *cap ? *cap * 2 : 8;
        *a = realloc(*a, sizeof(int64_t) * *cap);
    }
    (*a)[(*n)++] = v;
}

/* voronoi_edge_intersects_cocone (cocone.h:36-58) */
static int edge_intersects_cocone(double cos_a, double cos_b)
{
    if (fabs(cos_a) < COS_COCONE || fabs(cos_b) < COS_COCONE) return 1;
    if (cos_a < 0 && cos_b > 0) return 1;
    if (cos_a > 0 && cos_b < 0) return 1;
    return 0;
}

/* quadratic_equation (numerical/quadratic.h:25-48) */
static int quadratic(double a, double b, double c, double* r1, double* r2)
{
    const double d = b * b - 4 * a * c;
    if (d < 0) return 0;
    const double s = sqrt(d);
    if (b >= 0) {
        *r1 = (-b - s) / (2 * a);
        *r2 = 2 * c / (-b - s);
    } else {
        *r1 = (-b + s) / (2 * a);
        *r2 = 2 * c / (-b + s);
    }
    return 1;
}

/* intersect_cocone (cocone.h:84-157): the farthest point of the ray
 * a + t ab, t >= 0, on the double cone of half-angle 3 pi / 8 about the axis. */
static int intersect_cocone(int n, const double* axis, const double* a, const double* ab, double* distance)
{
    const double n_ab = dot_c(axis, ab, n);
    const double a_n = dot_c(a, axis, n);
    const double square_a = dot_c(a, a, n);
    const double square_ab = dot_c(ab, ab, n);
    const double a_ab = dot_c(a, ab, n);
    const double square_cos = COS_COCONE * COS_COCONE;
    const double qa = n_ab * n_ab - square_cos * square_ab;
    const double qb = 2 * (a_n * n_ab - a_ab * square_cos);
    const double qc = a_n * a_n - square_a * square_cos;
    double t1, t2;
    if (!quadratic(qa, qb, qc, &t1, &t2)) return 0;
    const int ok1 = t1 >= 0 && t1 <= DBL_MAX;
    const int ok2 = t2 >= 0 && t2 <= DBL_MAX;
    if (!ok1 && !ok2) return 0;
    double p1[CM], p2[CM];
    for (int i = 0; i < n; ++i) {
        p1[i] = a[i] + t1 * ab[i];
        p2[i] = a[i] + t2 * ab[i];
    }
    if (ok1 && !ok2) {
        *distance = length_c(p1, n);
        return 1;
    }
    if (!ok1 && ok2) {
        *distance = length_c(p2, n);
        return 1;
    }
    const double d1 = dot_c(p1, p1, n), d2 = dot_c(p2, p2, n);
    *distance = d1 > d2 ? sqrt(d1) : sqrt(d2);
    return 1;
}

static int inside_or_equal(double c) { return fabs(c) <= COS_COCONE; }

typedef struct {
    int n;
    const double* points;
    int64_t count;
    const int32_t* simplices;
    int64_t S;
    double* vv;  /* Voronoi vertex per simplex */
    cfacet* facets;
    int64_t F;
    cvertex* vtx;
    const char* err;
} cstate;

/* voronoi_edge_radius (structure.cpp:139-213) */
static double edge_radius(cstate* st, const cfacet* f, const double* pole, const double* pa, double pa_len,
                          double pb_len, double cos_a, double cos_b)
{
    const int n = st->n;
    const int one = f->s1 < 0;
    if (one && inside_or_equal(cos_b)) return DBL_MAX;
    if (!one && inside_or_equal(cos_a) && inside_or_equal(cos_b)) return pa_len > pb_len ? pa_len : pb_len;
    double ab[CM];
    for (int i = 0; i < n; ++i)
        ab[i] = one ? f->ortho[i] : st->vv[f->s1 * n + i] - st->vv[f->s0 * n + i];
    double maxd;
    if (!intersect_cocone(n, pole, pa, ab, &maxd)) {
        if (fabs(cos_a) > LIMIT_COSINE_FOR_INTERSECTION_PA_POLE) {
            const double ab_len = one ? 1.0 : length_c(ab, n);
            const double cos_pa_ab = dot_c(pa, ab, n) / (pa_len * ab_len);
            if (cos_pa_ab < LIMIT_COSINE_FOR_INTERSECTION_PA_AB) {
                maxd = 0;
            } else {
                st->err = "Cocone intersection not found, PA is close to positive pole";
                return -1;
            }
        } else {
            st->err = "Cocone intersection not found, PA is far from positive pole";
            return -1;
        }
    }
    if (!isfinite(maxd)) {
        st->err = "Cocone intersection distance is not finite";
        return -1;
    }
    if (inside_or_equal(cos_a)) return pa_len > maxd ? pa_len : maxd;
    return maxd;
}

/* vertex_and_facet_data (structure.cpp:343-398) with voronoi_positive_norm
 * (:61-104), voronoi_height (:106-137), cocone_facets_and_voronoi_radius
 * (:215-268) and cocone_neighbors (:270-302) */
static int vertex_data(cstate* st, int find_all)
{
    const int n = st->n;
    for (int64_t v = 0; v < st->count; ++v) {
        cvertex* c = &st->vtx[v];
        const double* p = st->points + v * n;
        memset(c->pn, 0, sizeof c->pn);
        c->height = 0;
        c->radius = 0;
        if (c->nf == 0 && c->no == 0) continue;  /* not a Delaunay vertex (a dropped duplicate) */
        /* positive pole */
        int unbounded = 0;
        for (int64_t k = 0; k < c->nf; ++k)
            if (st->facets[c->facet[k]].s1 < 0) unbounded = 1;
        if (unbounded) {
            double sum[CM] = {0};
            for (int64_t k = 0; k < c->nf; ++k) {
                const cfacet* f = &st->facets[c->facet[k]];
                if (f->s1 >= 0) continue;
                for (int i = 0; i < n; ++i) sum[i] = sum[i] + f->ortho[i];
            }
            memcpy(c->pn, sum, sizeof(double) * n);
            normalize_c(c->pn, n);
        } else {
            double maxd = -DBL_MAX, mv[CM] = {0};
            for (int64_t k = 0; k < c->no; ++k) {
                double vp[CM];
                for (int i = 0; i < n; ++i) vp[i] = st->vv[c->obj[k] * n + i] - p[i];
                const double d = dot_c(vp, vp, n);
                if (d > maxd) {
                    maxd = d;
                    memcpy(mv, vp, sizeof(double) * n);
                }
            }
            memcpy(c->pn, mv, sizeof(double) * n);
            normalize_c(c->pn, n);
        }
        if (!finite_c(c->pn, n)) {
            st->err = "Positive pole vector not finite";
            return 1;
        }
        if (find_all) {
            double maxd = -DBL_MAX;
            int found = 0;
            for (int64_t k = 0; k < c->no; ++k) {
                double vp[CM];
                for (int i = 0; i < n; ++i) vp[i] = st->vv[c->obj[k] * n + i] - p[i];
                if (dot_c(vp, c->pn, n) >= 0) continue;
                const double d = dot_c(vp, vp, n);
                if (d > maxd) {
                    maxd = d;
                    found = 1;
                }
            }
            if (!found) {
                st->err = "Negative pole vector not found";
                return 1;
            }
            c->height = sqrt(maxd);
            if (!isfinite(c->height)) {
                st->err = "Negative pole vector not finite";
                return 1;
            }
        }
        double radius = 0;
        for (int64_t k = 0; k < c->nf; ++k) {
            cfacet* f = &st->facets[c->facet[k]];
            double pa[CM];
            for (int i = 0; i < n; ++i) pa[i] = st->vv[f->s0 * n + i] - p[i];
            const double pa_len = length_c(pa, n);
            const double cos_a = dot_c(c->pn, pa, n) / pa_len;
            double pb_len, cos_b;
            if (f->s1 < 0) {
                pb_len = 0;
                cos_b = dot_c(c->pn, f->ortho, n);
            } else {
                double pb[CM];
                for (int i = 0; i < n; ++i) pb[i] = st->vv[f->s1 * n + i] - p[i];
                pb_len = length_c(pb, n);
                cos_b = dot_c(c->pn, pb, n) / pb_len;
            }
            if (!edge_intersects_cocone(cos_a, cos_b)) continue;
            f->cocone[c->local[k]] = 1;
            if (find_all && radius != DBL_MAX) {
                const double r = edge_radius(st, f, c->pn, pa, pa_len, pb_len, cos_a, cos_b);
                if (r < 0) return 1;
                if (r > radius) radius = r;
            }
        }
        if (find_all && !(radius > 0 && radius <= DBL_MAX)) {
            st->err = "Voronoi cell radius is not positive";
            return 1;
        }
        c->radius = radius;
    }
    if (find_all) {
        for (int64_t v = 0; v < st->count; ++v) {
            cvertex* c = &st->vtx[v];
            for (int64_t k = 0; k < c->nf; ++k) {
                const cfacet* f = &st->facets[c->facet[k]];
                for (int u = 0; u < n; ++u) {
                    if (u == c->local[k]) continue;
                    if (f->cocone[u]) {
                        if (c->nn == c->cn) {
                            c->cn = c->cn ? c->cn * 2 : 8;
                            c->nbr = realloc(c->nbr, sizeof(int32_t) * c->cn);
                        }
                        c->nbr[c->nn++] = f->v[u];
                    }
                }
            }
            /* sort_and_unique */
            for (int64_t i = 1; i < c->nn; ++i) {
                int32_t t = c->nbr[i];
                int64_t j = i - 1;
                while (j >= 0 && c->nbr[j] > t) {
                    c->nbr[j + 1] = c->nbr[j];
                    --j;
                }
                c->nbr[j + 1] = t;
            }
            int64_t w = 0;
            for (int64_t i = 0; i < c->nn; ++i)
                if (w == 0 || c->nbr[w - 1] != c->nbr[i]) c->nbr[w++] = c->nbr[i];
            c->nn = w;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* prune_facets_incident_to_sharp_ridges (prune_facets.h:117-200)             */

typedef struct {
    int32_t v[CM];       /* the ridge's n - 1 sorted vertices */
    int64_t facet;       /* a facet holding it */
    int32_t point;       /* the facet's vertex outside the ridge */
    int64_t order;       /* insertion order within the ridge's list */
    int alive;
} ridge_rec;

static int g_k;

static int cmp_ridge_v(const int32_t* a, const int32_t* b)
{
    for (int i = 0; i < g_k; ++i)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}

static int cmp_ridge_rec(const void* a, const void* b)
{
    const ridge_rec* x = a;
    const ridge_rec* y = b;
    const int c = cmp_ridge_v(x->v, y->v);
    if (c) return c;
    return x->order < y->order ? -1 : x->order > y->order;
}

/* sharp_ridge (prune_facets.h:56-115) for the ridge's live facets e[0..m) in
 * insertion order; ortho_e0_e1 (linear_algebra.h:332-352) */
static int sharp_ridge(cstate* st, const int* interior, const ridge_rec* const* e, int m)
{
    const int n = st->n;
    const int32_t* rv = e[0]->v;
    for (int i = 0; i < n - 1; ++i)
        if (!interior[rv[i]]) return 0;  /* boundary ridge */
    if (m == 1) return 1;
    const double* P = st->points;
    double vecs[CM * CM], e0[CM], e1[CM];
    for (int i = 0; i < n - 2; ++i)
        for (int c = 0; c < n; ++c) vecs[i * n + c] = P[(int64_t)rv[i + 1] * n + c] - P[(int64_t)rv[0] * n + c];
    for (int c = 0; c < n; ++c) vecs[(n - 2) * n + c] = P[(int64_t)e[0]->point * n + c] - P[(int64_t)rv[0] * n + c];
    orc_ortho_nn(n, vecs, e0);
    normalize_c(e0, n);
    memcpy(vecs + (n - 2) * n, e0, sizeof(double) * n);
    orc_ortho_nn(n, vecs, e1);
    normalize_c(e1, n);
    double bv[CM];
    for (int c = 0; c < n; ++c) bv[c] = P[(int64_t)e[0]->point * n + c] - P[(int64_t)rv[0] * n + c];
    double base[2] = {dot_c(e0, bv, n), dot_c(e1, bv, n)};
    normalize_c(base, 2);
    double cos_plus = 1, cos_minus = 1, sin_plus = 0, sin_minus = 0;
    for (int k = 1; k < m; ++k) {
        double fv[CM];
        for (int c = 0; c < n; ++c) fv[c] = P[(int64_t)e[k]->point * n + c] - P[(int64_t)rv[0] * n + c];
        double w[2] = {dot_c(e0, fv, n), dot_c(e1, fv, n)};
        normalize_c(w, 2);
        const double sine = base[0] * w[1] - base[1] * w[0];
        const double cosine = dot_c(base, w, 2);
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
    if (cos_plus <= 0 || cos_minus <= 0) return 0;
    const double cos_a_plus_b = cos_plus * cos_minus - fabs(sin_plus * sin_minus);
    return cos_a_plus_b > 0;
}

static int64_t ridge_find(const ridge_rec* R, int64_t nr, const int32_t* v)
{
    int64_t lo = 0, hi = nr;
    while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        if (cmp_ridge_v(R[mid].v, v) < 0) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

static int cmp_vkey(const void* a, const void* b) { return cmp_ridge_v(a, b); }

static int prune(cstate* st, const int* interior, char* keep)
{
    const int n = st->n;
    g_k = n - 1;
    /* every (ridge, facet) pair of the kept facets, in facet order */
    int64_t nr = 0;
    for (int64_t f = 0; f < st->F; ++f)
        if (keep[f]) nr += n;
    ridge_rec* R = malloc(sizeof(ridge_rec) * (nr ? nr : 1));
    int64_t w = 0;
    for (int64_t f = 0; f < st->F; ++f) {
        if (!keep[f]) continue;
        for (int r = 0; r < n; ++r) {
            ridge_rec* x = &R[w];
            int c = 0;
            for (int i = 0; i < n; ++i)
                if (i != r) x->v[c++] = st->facets[f].v[i];
            x->facet = f;
            x->point = st->facets[f].v[r];
            x->order = w;
            x->alive = 1;
            ++w;
        }
    }
    qsort(R, nr, sizeof(ridge_rec), cmp_ridge_rec);
    /* suspicious ridges: all, then the other ridges of removed facets */
    int32_t* sus = malloc(sizeof(int32_t) * CM * (nr ? nr : 1));
    int64_t ns = 0;
    for (int64_t i = 0; i < nr; ++i)
        if (i == 0 || cmp_ridge_v(R[i].v, R[i - 1].v)) memcpy(sus + (ns++) * CM, R[i].v, sizeof(int32_t) * CM);
    int32_t* nxt = malloc(sizeof(int32_t) * CM * (nr ? nr : 1));
    const ridge_rec** e = malloc(sizeof(ridge_rec*) * (nr ? nr : 1));
    while (ns > 0) {
        int64_t nn = 0;
        for (int64_t s = 0; s < ns; ++s) {
            const int32_t* rv = sus + s * CM;
            int64_t i = ridge_find(R, nr, rv);
            int m = 0;
            for (; i < nr && !cmp_ridge_v(R[i].v, rv); ++i)
                if (R[i].alive) e[m++] = &R[i];
            if (m == 0) continue;  /* no longer in the map */
            if (!sharp_ridge(st, interior, e, m)) continue;
            for (int k = 0; k < m; ++k) {
                const int64_t f = e[k]->facet;
                const int32_t point = e[k]->point;
                /* the facet's ridges that contain point: all but this one */
                for (int r = 0; r < n; ++r) {
                    if (st->facets[f].v[r] == point) continue;
                    int32_t* x = nxt + (nn++) * CM;
                    memset(x, 0, sizeof(int32_t) * CM);
                    int c = 0;
                    for (int q = 0; q < n; ++q)
                        if (q != r) x[c++] = st->facets[f].v[q];
                }
                keep[f] = 0;
            }
            /* remove the facets from every ridge they hold */
            for (int k = 0; k < m; ++k) {
                const int64_t f = e[k]->facet;
                for (int r = 0; r < n; ++r) {
                    int32_t rv2[CM] = {0};
                    int c = 0;
                    for (int q = 0; q < n; ++q)
                        if (q != r) rv2[c++] = st->facets[f].v[q];
                    for (int64_t j = ridge_find(R, nr, rv2); j < nr && !cmp_ridge_v(R[j].v, rv2); ++j)
                        if (R[j].facet == f) R[j].alive = 0;
                }
            }
        }
        /* the next round's ridges, sorted and unique */
        qsort(nxt, nn, sizeof(int32_t) * CM, cmp_vkey);
        ns = 0;
        for (int64_t i = 0; i < nn; ++i)
            if (i == 0 || cmp_ridge_v(nxt + i * CM, nxt + (i - 1) * CM)) memcpy(sus + (ns++) * CM, nxt + i * CM, sizeof(int32_t) * CM);
    }
    free(R);
    free(sus);
    free(nxt);
    free(e);
    return 0;
}

/* extract_manifold (extract_manifold.h:105-131) with traverse_delaunay
 * (:50-103): the Delaunay objects reachable from the hull's one-sided facets
 * without crossing a kept facet; the kept facets met on the way remain. */
static void extract(cstate* st, char* keep)
{
    const int n = st->n;
    int64_t* start = calloc(st->S + 1, sizeof(int64_t));
    for (int64_t f = 0; f < st->F; ++f) {
        start[st->facets[f].s0 + 1]++;
        if (st->facets[f].s1 >= 0) start[st->facets[f].s1 + 1]++;
    }
    for (int64_t s = 0; s < st->S; ++s) start[s + 1] += start[s];
    int64_t* of = malloc(sizeof(int64_t) * (start[st->S] ? start[st->S] : 1));
    int64_t* fill = calloc(st->S, sizeof(int64_t));
    for (int64_t f = 0; f < st->F; ++f) {
        of[start[st->facets[f].s0] + fill[st->facets[f].s0]++] = f;
        if (st->facets[f].s1 >= 0) of[start[st->facets[f].s1] + fill[st->facets[f].s1]++] = f;
    }
    char* vis = calloc(st->S, 1);
    char* met = calloc(st->F, 1);
    int64_t cap = st->F + 16, top = 0;
    int64_t* stack = malloc(sizeof(int64_t) * cap);
    for (int64_t f = 0; f < st->F; ++f)
        if (st->facets[f].s1 < 0) stack[top++] = f;
    while (top > 0) {
        const int64_t f = stack[--top];
        if (keep[f]) {
            met[f] = 1;
            continue;
        }
        const cfacet* fc = &st->facets[f];
        int64_t d;
        if (fc->s1 < 0) {
            if (vis[fc->s0]) continue;
            d = fc->s0;
        } else {
            if (vis[fc->s0] && vis[fc->s1]) continue;
            d = vis[fc->s0] ? fc->s1 : fc->s0;
        }
        vis[d] = 1;
        for (int64_t k = start[d]; k < start[d + 1]; ++k) {
            if (of[k] == f) continue;
            if (top == cap) {
                cap *= 2;
                stack = realloc(stack, sizeof(int64_t) * cap);
            }
            stack[top++] = of[k];
        }
    }
    memcpy(keep, met, st->F);
    (void)n;
    free(start);
    free(of);
    free(fill);
    free(vis);
    free(met);
    free(stack);
}

static void free_state(cstate* st)
{
    if (st->vtx) {
        for (int64_t v = 0; v < st->count; ++v) {
            free(st->vtx[v].facet);
            free(st->vtx[v].local);
            free(st->vtx[v].obj);
            free(st->vtx[v].nbr);
        }
    }
    free(st->vtx);
    free(st->facets);
    free(st->vv);
}

int64_t orc_cocone(int n, const double* points, int64_t count, const int32_t* simplices, const double* orthos,
                   int64_t S, int mode, double rho, double alpha, int32_t* facets_out, int64_t facets_cap,
                   double* normals_out)
{
    if (n < 2 || n > 5) {
        orc_set_error("cocone dimension must be in [2, 5]");
        return -1;
    }
    cstate st;
    memset(&st, 0, sizeof st);
    st.n = n;
    st.points = points;
    st.count = count;
    st.simplices = simplices;
    st.S = S;
    if (mode == 1 && !(rho > 0 && rho < 1)) {
        orc_set_error("Rho must be in the interval (0, 1)");
        return -1;
    }
    if (mode == 1 && !(alpha > 0 && alpha < 1)) {
        orc_set_error("Alpha must be in the interval (0, 1)");
        return -1;
    }
    /* Voronoi vertices */
    st.vv = malloc(sizeof(double) * n * (S ? S : 1));
    for (int64_t s = 0; s < S; ++s) {
        orc_voronoi_vertex(n, points, simplices + s * (n + 1), st.vv + s * n);
        if (!finite_c(st.vv + s * n, n)) {
            orc_set_error("Voronoi vertex is not finite");
            free_state(&st);
            return -1;
        }
    }
    /* facets: every face of every simplex, grouped by sorted vertex tuple */
    g_n = n;
    face_rec* fr = malloc(sizeof(face_rec) * (S ? S : 1) * (n + 1));
    int64_t nf = 0;
    for (int64_t s = 0; s < S; ++s) {
        for (int r = 0; r <= n; ++r) {
            face_rec* x = &fr[nf++];
            memset(x->v, 0, sizeof x->v);
            int c = 0;
            for (int i = 0; i <= n; ++i)
                if (i != r) x->v[c++] = simplices[s * (n + 1) + i];
            sort_ints(x->v, n);
            x->s = s;
            x->r = r;
        }
    }
    qsort(fr, nf, sizeof(face_rec), cmp_face);
    st.facets = malloc(sizeof(cfacet) * (nf ? nf : 1));
    for (int64_t i = 0; i < nf;) {
        int64_t j = i + 1;
        while (j < nf && !memcmp(fr[j].v, fr[i].v, sizeof(int32_t) * n)) ++j;
        if (j - i > 2) {
            orc_set_error("a Delaunay facet in more than two simplices");
            free(fr);
            free_state(&st);
            return -1;
        }
        cfacet* f = &st.facets[st.F++];
        memset(f, 0, sizeof *f);
        memcpy(f->v, fr[i].v, sizeof(int32_t) * n);
        f->s0 = fr[i].s;
        f->s1 = j - i == 2 ? fr[i + 1].s : -1;
        memcpy(f->ortho, orthos + (fr[i].s * (n + 1) + fr[i].r) * n, sizeof(double) * n);
        i = j;
    }
    free(fr);
    /* vertex connections (structure.cpp:304-325) */
    st.vtx = calloc(count, sizeof(cvertex));
    for (int64_t f = 0; f < st.F; ++f) {
        for (int k = 0; k < n; ++k) {
            cvertex* c = &st.vtx[st.facets[f].v[k]];
            if (c->nf == c->cf) {
                c->cf = c->cf ? c->cf * 2 : 8;
                c->facet = realloc(c->facet, sizeof(int64_t) * c->cf);
                c->local = realloc(c->local, sizeof(int) * c->cf);
            }
            c->facet[c->nf] = f;
            c->local[c->nf] = k;
            ++c->nf;
        }
    }
    for (int64_t s = 0; s < S; ++s)
        for (int k = 0; k <= n; ++k) {
            cvertex* c = &st.vtx[simplices[s * (n + 1) + k]];
            push_i64(&c->obj, &c->no, &c->co, s);
        }
    if (vertex_data(&st, mode == 1)) {
        orc_set_error(st.err);
        free_state(&st);
        return -1;
    }
    char* keep = calloc(st.F ? st.F : 1, 1);
    int* interior = malloc(sizeof(int) * (count ? count : 1));
    if (mode == 0) {
        /* cocone facets: the Voronoi edge meets the cocone of all n vertices */
        int any = 0;
        for (int64_t f = 0; f < st.F; ++f) {
            int all = 1;
            for (int k = 0; k < n; ++k) all = all && st.facets[f].cocone[k];
            keep[f] = (char)all;
            any |= all;
        }
        if (!any) {
            orc_set_error("Cocone facets not found");
            goto fail;
        }
        for (int64_t v = 0; v < count; ++v) interior[v] = 1;
    } else {
        /* find_interior_vertices (reconstruction.cpp:72-160) */
        const double cos_alpha = cos(alpha);
        int any = 0;
        for (int64_t v = 0; v < count; ++v) {
            const cvertex* c = &st.vtx[v];
            interior[v] = 0;
            if (!(c->radius <= rho * c->height)) continue;
            int flat = 1;
            for (int64_t k = 0; k < c->nn && flat; ++k)
                if (!(fabs(dot_c(c->pn, st.vtx[c->nbr[k]].pn, n)) >= cos_alpha)) flat = 0;
            if (flat) {
                interior[v] = 1;
                any = 1;
            }
        }
        if (any) {
            int found;
            do {
                found = 0;
                for (int64_t v = 0; v < count; ++v) {
                    if (interior[v]) continue;
                    const cvertex* c = &st.vtx[v];
                    if (!(c->radius <= rho * c->height)) continue;
                    for (int64_t k = 0; k < c->nn; ++k) {
                        if (!interior[c->nbr[k]]) continue;
                        if (fabs(dot_c(c->pn, st.vtx[c->nbr[k]].pn, n)) >= cos_alpha) {
                            interior[v] = 1;
                            found = 1;
                            break;
                        }
                    }
                }
            } while (found);
        }
        if (!any) {
            orc_set_error("Interior vertices not found");
            goto fail;
        }
        /* cocone_interior_facet (reconstruction.cpp:162-181) */
        int anyf = 0;
        for (int64_t f = 0; f < st.F; ++f) {
            int ok = 1, fnd = 0;
            for (int k = 0; k < n && ok; ++k) {
                const int in = interior[st.facets[f].v[k]];
                const int ic = in && st.facets[f].cocone[k];
                if (!(ic || !in)) ok = 0;
                fnd = fnd || ic;
            }
            keep[f] = (char)(ok && fnd);
            anyf |= keep[f];
        }
        if (!anyf) {
            orc_set_error("Cocone interior facets not found");
            goto fail;
        }
    }
    prune(&st, interior, keep);
    {
        int any = 0;
        for (int64_t f = 0; f < st.F; ++f) any |= keep[f];
        if (!any) {
            orc_set_error("Cocone facets not found after prune");
            goto fail;
        }
    }
    extract(&st, keep);
    int64_t out = 0;
    for (int64_t f = 0; f < st.F; ++f) {
        if (!keep[f]) continue;
        if (facets_out && out < facets_cap) memcpy(facets_out + out * n, st.facets[f].v, sizeof(int32_t) * n);
        ++out;
    }
    if (out == 0) {
        orc_set_error("Cocone facets not found after manifold extraction");
        goto fail;
    }
    if (normals_out) {
        memset(normals_out, 0, sizeof(double) * n * count);
        for (int64_t f = 0; f < st.F; ++f) {
            if (!keep[f]) continue;
            for (int k = 0; k < n; ++k) {
                const int32_t p = st.facets[f].v[k];
                memcpy(normals_out + (int64_t)p * n, st.vtx[p].pn, sizeof(double) * n);
            }
        }
    }
    free(keep);
    free(interior);
    free_state(&st);
    return out;
fail:
    free(keep);
    free(interior);
    free_state(&st);
    return -1;
}
