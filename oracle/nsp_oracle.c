/*
 * nsp_oracle.c -- FP64 CPU restatement of the reference painter path.
 * TEST INFRASTRUCTURE ONLY (see nsp_oracle.h).  Citations are relative to the
 * templeblock/math source tree.
 */
#define _GNU_SOURCE
#include "nsp_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NM ORC_NMAX

static __thread char g_err[512];

const char* orc_last_error(void) { return g_err; }

static int fail(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return 1;
}

void orc_set_error(const char* msg) { fail("%s", msg); }

/* ------------------------------------------------------------------------ */
/* Philox4x32-10.  Constants from Salmon et al. (SC'11), Random123.           */

void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static double u01(uint32_t x) { return (double)(x >> 8) * (1.0 / 16777216.0); }

/* Four uniforms of one stream: counter (pixel lo, sample, pass|depth|stream, pixel hi). */
static void uniforms4(uint64_t seed, int64_t pixel, int sample, int pass, int depth, int stream, double u[4])
{
    uint32_t ctr[4] = {(uint32_t)pixel, (uint32_t)sample,
                       ((uint32_t)pass << 16) | ((uint32_t)depth << 8) | (uint32_t)stream,
                       (uint32_t)((uint64_t)pixel >> 32)};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    orc_philox4x32(ctr, key, o);
    for (int i = 0; i < 4; ++i) u[i] = u01(o[i]);
}

/* ------------------------------------------------------------------------ */
/* Vector helpers.  dot is the fma chain of src/com/vec.h:258-267.            */

static double dotn(const double* a, const double* b, int n)
{
    double r = a[0] * b[0];
    for (int i = 1; i < n; ++i) r = fma(a[i], b[i], r);
    return r;
}

static void normalize_n(double* v, int n)
{
    double len = sqrt(dotn(v, v, n));
    for (int i = 0; i < n; ++i) v[i] = v[i] / len;
}

static int finite_n(const double* v, int n)
{
    for (int i = 0; i < n; ++i)
        if (!isfinite(v[i])) return 0;
    return 1;
}

/* Determinant by cofactor expansion along the first row of the selected rows
 * and columns, src/geometry/core/linear_algebra.h:60-104 (the explicit N=3,4
 * forms at :194-236 are this expansion written out). */
static double det_cof(const double* m, int stride, const int* rows, const int* cols, int k)
{
    if (k == 1) return m[rows[0] * stride + cols[0]];
    if (k == 2) {
        return m[rows[0] * stride + cols[0]] * m[rows[1] * stride + cols[1]] -
               m[rows[0] * stride + cols[1]] * m[rows[1] * stride + cols[0]];
    }
    double s = 0;
    int sub[NM];
    for (int j = 0; j < k; ++j) {
        int c = 0;
        for (int q = 0; q < k; ++q)
            if (q != j) sub[c++] = cols[q];
        double term = m[rows[0] * stride + cols[j]] * det_cof(m, stride, rows + 1, sub, k - 1);
        if (j == 0) s = term;
        else if (j & 1) s = s - term;
        else s = s + term;
    }
    return s;
}

/* ortho_nn, src/geometry/core/linear_algebra.h:167-186 (general), :194-236 (N=3,4). */
void orc_ortho_nn(int n, const double* vectors, double* out)
{
    if (n == 2) {
        out[0] = vectors[1];
        out[1] = -vectors[0];
        return;
    }
    int rows[NM], cols[NM];
    for (int r = 0; r < n - 1; ++r) rows[r] = r;
    for (int i = 0; i < n; ++i) {
        int c = 0;
        for (int q = 0; q < n; ++q)
            if (q != i) cols[c++] = q;
        double minor = det_cof(vectors, n, rows, cols, n - 1);
        out[i] = (i & 1) ? -minor : minor;
    }
}

/* cross, src/com/vec.h:417-431 */
static void cross3(const double* a, const double* b, double* r)
{
    r[0] = +(a[1] * b[2] - a[2] * b[1]);
    r[1] = -(a[0] * b[2] - a[2] * b[0]);
    r[2] = +(a[0] * b[1] - a[1] * b[0]);
}

/* orthogonal_complement_of_unit_vector, src/geometry/core/complement.h:71-200 */
void orc_orthogonal_complement(int n, const double* u, double* basis)
{
    const double LIMIT = 0.1;
    if (n == 2) {
        basis[0] = u[1];
        basis[1] = -u[0];
        return;
    }
    if (n == 3) {
        double nc[3] = {0, 0, 0};
        if (fabs(u[0]) > LIMIT) nc[1] = 1; else nc[0] = 1;
        double e0[3], e1[3];
        cross3(u, nc, e0);
        normalize_n(e0, 3);
        cross3(u, e0, e1);
        memcpy(basis, e0, sizeof e0);
        memcpy(basis + 3, e1, sizeof e1);
        return;
    }
    if (n == 4) {
        /* by subspace, complement.h:71-122 */
        int exclude = 0;
        for (; exclude < n - 2; ++exclude)
            if (fabs(u[exclude]) > LIMIT) break;
        double sb[NM * NM];
        memset(sb, 0, sizeof sb);
        memcpy(sb + (n - 2) * n, u, n * sizeof(double));
        int num = 0;
        for (int i = 0; num < n - 2; ++i)
            if (i != exclude) {
                for (int q = 0; q < n; ++q) sb[num * n + q] = (q == i) ? 1.0 : 0.0;
                ++num;
            }
        double t[NM];
        for (int i = 0; i < n - 2; ++i) {
            orc_ortho_nn(n, sb, t);
            normalize_n(t, n);
            memcpy(sb + i * n, t, n * sizeof(double));
        }
        orc_ortho_nn(n, sb, t);
        memcpy(sb + (n - 2) * n, t, n * sizeof(double));
        memcpy(basis, sb, (n - 1) * n * sizeof(double));
        return;
    }
    /* by Gram-Schmidt, complement.h:124-187 */
    int exclude = 0;
    for (; exclude < n - 1; ++exclude)
        if (fabs(u[exclude]) > LIMIT) break;
    double b[NM][NM], ob[NM][NM];
    memcpy(b[0], u, n * sizeof(double));
    int num = 1;
    for (int i = 0; num < n; ++i)
        if (i != exclude) {
            for (int q = 0; q < n; ++q) b[num][q] = (q == i) ? 1.0 : 0.0;
            ++num;
        }
    memcpy(ob[0], b[0], n * sizeof(double));
    for (int i = 1; i < n; ++i) {
        double sum[NM] = {0};
        for (int k = 0; k < i; ++k) {
            double d = dotn(b[i], ob[k], n);
            for (int q = 0; q < n; ++q) sum[q] = sum[q] + d * ob[k][q];
        }
        for (int q = 0; q < n; ++q) ob[i][q] = b[i][q] - sum[q];
        normalize_n(ob[i], n);
    }
    for (int i = 0; i < n - 1; ++i) memcpy(basis + i * n, ob[i + 1], n * sizeof(double));
}

/* cosine_sphere_coefficient, src/painter/coefficient/cosine_sphere.h:97-162 */
double orc_cosine_sphere_coefficient(int n)
{
    /* (n-1)(n-3)... / (n-2)(n-4)..., times pi/2 for even n */
    long double p = (n & 1) == 0 ? 0.5L : 1.0L;
    for (int i = n - 1; i > 1; i -= 2) {
        p *= i;
        if (i > 2) p /= (i - 1);
    }
    if ((n & 1) == 0) p *= 3.14159265358979323846264338327950288L;
    return (double)p;
}

/* ------------------------------------------------------------------------ */
/* Mesh                                                                       */

typedef struct {
    int level;
    int64_t q[NM];
    int child0; /* index of first child in nodes[], -1 for a leaf */
    int64_t begin, count; /* leaf entries */
} onode;

struct orc_mesh {
    int n;
    int64_t V, F, VN;
    double* vtx;   /* V*n */
    double* vn;    /* VN*n (may be NULL) */
    int32_t* fac;  /* F*n */
    int32_t* fnrm; /* F*n or NULL */
    double* normal;  /* F*n */
    double* pn;      /* F*(n-1)*n */
    double* pd;      /* F*(n-1) */
    char* ntype;     /* 0 None, 1 Use, 2 Reverse */
    uint8_t* rev;    /* F*n */
    double mn[NM], mx[NM];
    /* tree */
    int has_tree;
    double root_lo[NM], root_size[NM];
    double dist_from_facet;
    int max_depth;
    onode* nodes;
    int64_t n_nodes, cap_nodes;
    int32_t* entries;
    int64_t n_entries, cap_entries;
    double* cons;  /* F * n constraints (n+1 each: a[n], b) */
    double* ceq;   /* F * (n+1) */
    double* smin;  /* F*n */
    double* smax;
};

/* HyperplaneSimplexGeometry::set_data, src/painter/space/hyperplane_geometry.h:83-115 */
static void set_data(int n, const double* normal, const double* verts /* n*n */, double* pn, double* pd)
{
    double vectors[NM * NM];
    for (int i = 0; i < n - 1; ++i)
        for (int q = 0; q < n; ++q) vectors[i * n + q] = verts[i * n + q] - verts[(n - 1) * n + q];
    for (int i = 0; i < n - 1; ++i) {
        double saved[NM];
        memcpy(saved, vectors + i * n, n * sizeof(double));
        memcpy(vectors + i * n, normal, n * sizeof(double));
        double* pni = pn + i * n;
        orc_ortho_nn(n, vectors, pni);
        memcpy(vectors + i * n, saved, n * sizeof(double));
        double d = dotn(verts + (n - 1) * n, pni, n);
        double distance = dotn(verts + i * n, pni, n) - d;
        for (int q = 0; q < n; ++q) pni[q] = pni[q] / distance;
        pd[i] = d / distance;
    }
}

/* HyperplaneSimplexGeometry::constraints, hyperplane_geometry.h:120-160 */
static void simplex_constraints(int n, const double* normal, const double* verts, const double* pn,
                                const double* pd, double* c /* n*(n+1) */, double* ceq /* n+1 */)
{
    for (int i = 0; i < n - 1; ++i) {
        double len = sqrt(dotn(pn + i * n, pn + i * n, n));
        for (int q = 0; q < n; ++q) c[i * (n + 1) + q] = pn[i * n + q] / len;
        c[i * (n + 1) + n] = -pd[i] / len;
    }
    double vectors[NM * NM];
    for (int i = 0; i < n - 2; ++i)
        for (int q = 0; q < n; ++q) vectors[i * n + q] = verts[(i + 1) * n + q] - verts[q];
    memcpy(vectors + (n - 2) * n, normal, n * sizeof(double));
    double nn[NM];
    orc_ortho_nn(n, vectors, nn);
    normalize_n(nn, n);
    double d = dotn(verts, nn, n);
    int to_vertex = dotn(verts + (n - 1) * n, nn, n) - d >= 0;
    for (int q = 0; q < n; ++q) c[(n - 1) * (n + 1) + q] = to_vertex ? nn[q] : -nn[q];
    c[(n - 1) * (n + 1) + n] = to_vertex ? -d : d;
    d = dotn(verts, normal, n);
    memcpy(ceq, normal, n * sizeof(double));
    ceq[n] = -d;
}

int orc_mesh_create(int n, const float* vertices, int64_t V, const int32_t* facets, int64_t F,
                    const float* normals, int64_t VN, const int32_t* facet_normals, const double* matrix,
                    orc_mesh** out)
{
    *out = NULL;
    if (n < 3 || n > NM) return fail("dimension %d is not in [3, %d]", n, NM);
    if (V <= 0) return fail("No vertices found in obj");
    if (F <= 0) return fail("No facets found in obj");
    if (matrix) {
        /* MatrixMulVector, src/com/matrix_alg.h:30-44 */
        if (matrix[n * (n + 1) + n] != 1) return fail("Wrong matrix for matrix-vector multiplier");
        for (int i = 0; i < n; ++i)
            if (matrix[n * (n + 1) + i] != 0) return fail("Wrong matrix for matrix-vector multiplier");
    }
    orc_mesh* m = calloc(1, sizeof *m);
    m->n = n; m->V = V; m->F = F; m->VN = normals ? VN : 0;
    m->vtx = malloc(sizeof(double) * V * n);
    m->fac = malloc(sizeof(int32_t) * F * n);
    memcpy(m->fac, facets, sizeof(int32_t) * F * n);
    for (int64_t v = 0; v < V; ++v) {
        for (int i = 0; i < n; ++i) {
            if (!matrix) {
                m->vtx[v * n + i] = vertices[v * n + i];
            } else {
                /* src/com/matrix_alg.h:46-63 */
                double r = matrix[i * (n + 1) + n];
                for (int c = 0; c < n; ++c) r = fma(matrix[i * (n + 1) + c], (double)vertices[v * n + c], r);
                m->vtx[v * n + i] = r;
            }
        }
    }
    if (normals && facet_normals) {
        m->vn = malloc(sizeof(double) * VN * n);
        for (int64_t i = 0; i < VN * n; ++i) m->vn[i] = normals[i];
        m->fnrm = malloc(sizeof(int32_t) * F * n);
        memcpy(m->fnrm, facet_normals, sizeof(int32_t) * F * n);
    }
    m->normal = malloc(sizeof(double) * F * n);
    m->pn = malloc(sizeof(double) * F * (n - 1) * n);
    m->pd = malloc(sizeof(double) * F * (n - 1));
    m->ntype = calloc(F, 1);
    m->rev = calloc(F * n, 1);
    for (int i = 0; i < n; ++i) { m->mn[i] = INFINITY; m->mx[i] = -INFINITY; }

    for (int64_t f = 0; f < F; ++f) {
        const int32_t* v = m->fac + f * n;
        for (int k = 0; k < n; ++k)
            if (v[k] < 0 || v[k] >= V) {
                orc_mesh_destroy(m);
                return fail("facet %lld vertex index %d out of range", (long long)f, v[k]);
            }
        /* MeshHyperplaneSimplex ctor, src/painter/shapes/mesh_hyperplane_simplex.cpp:59-143 */
        double vec[NM * NM];
        for (int i = 0; i < n - 1; ++i)
            for (int q = 0; q < n; ++q) vec[i * n + q] = m->vtx[(int64_t)v[i + 1] * n + q] - m->vtx[(int64_t)v[0] * n + q];
        double* nrm = m->normal + f * n;
        orc_ortho_nn(n, vec, nrm);
        normalize_n(nrm, n);
        if (!finite_n(nrm, n)) {
            orc_mesh_destroy(m);
            return fail("Simplex normal is not finite, simplex %lld", (long long)f);
        }
        double verts[NM * NM];
        for (int k = 0; k < n; ++k) memcpy(verts + k * n, m->vtx + (int64_t)v[k] * n, n * sizeof(double));
        set_data(n, nrm, verts, m->pn + f * (n - 1) * n, m->pd + f * (n - 1));
        for (int k = 0; k < n; ++k)
            for (int i = 0; i < n; ++i) {
                m->mn[i] = fmin(m->mn[i], verts[k * n + i]);
                m->mx[i] = fmax(m->mx[i], verts[k * n + i]);
            }
        if (!m->vn) continue;
        const int32_t* nidx = m->fnrm + f * n;
        int has = 1;
        for (int k = 0; k < n; ++k)
            if (nidx[k] < 0 || nidx[k] >= VN) has = 0;
        if (!has) continue;
        double dots[NM];
        int any_small = 0, all_pos = 1, all_neg = 1;
        for (int k = 0; k < n; ++k) {
            dots[k] = dotn(m->vn + (int64_t)nidx[k] * n, nrm, n);
            if (fabs(dots[k]) < 0.7) any_small = 1;
            if (!(dots[k] > 0)) all_pos = 0;
            if (!(dots[k] < 0)) all_neg = 0;
        }
        if (any_small) continue;
        if (all_pos) { m->ntype[f] = 1; continue; }
        if (all_neg) {
            m->ntype[f] = 1;
            for (int i = 0; i < n; ++i) nrm[i] = -nrm[i];
            continue;
        }
        m->ntype[f] = 2;
        for (int k = 0; k < n; ++k) m->rev[f * n + k] = dots[k] < 0;
    }
    *out = m;
    return 0;
}

void orc_mesh_destroy(orc_mesh* m)
{
    if (!m) return;
    free(m->vtx); free(m->vn); free(m->fac); free(m->fnrm);
    free(m->normal); free(m->pn); free(m->pd); free(m->ntype); free(m->rev);
    free(m->nodes); free(m->entries); free(m->cons); free(m->ceq); free(m->smin); free(m->smax);
    free(m);
}

void orc_mesh_records(const orc_mesh* m, double* normal, double* plane_n, double* plane_d, double* vertex0)
{
    int n = m->n;
    if (normal) memcpy(normal, m->normal, sizeof(double) * m->F * n);
    if (plane_n) memcpy(plane_n, m->pn, sizeof(double) * m->F * (n - 1) * n);
    if (plane_d) memcpy(plane_d, m->pd, sizeof(double) * m->F * (n - 1));
    if (vertex0)
        for (int64_t f = 0; f < m->F; ++f)
            memcpy(vertex0 + f * n, m->vtx + (int64_t)m->fac[f * n] * n, n * sizeof(double));
}

void orc_mesh_min_max(const orc_mesh* m, double* mn, double* mx)
{
    memcpy(mn, m->mn, sizeof(double) * m->n);
    memcpy(mx, m->mx, sizeof(double) * m->n);
}

/* ------------------------------------------------------------------------ */
/* Ray / simplex                                                              */

/* HyperplaneSimplexGeometry::intersect, hyperplane_geometry.h:181-204 with
 * plane_intersect :40-52.  Returns 1 on hit; *margin = min barycentric coordinate. */
static int simplex_hit(const orc_mesh* m, int64_t f, const double* org, const double* dir, double* t,
                       double* margin)
{
    int n = m->n;
    const double* nrm = m->normal + f * n;
    const double* v0 = m->vtx + (int64_t)m->fac[f * n] * n;
    double s = dotn(nrm, dir, n);
    *margin = -INFINITY;
    if (s == 0) return 0;
    double diff[NM];
    for (int i = 0; i < n; ++i) diff[i] = v0[i] - org[i];
    *t = dotn(diff, nrm, n) / s;
    if (!(*t > 0)) return 0;
    double p[NM];
    for (int i = 0; i < n; ++i) p[i] = org[i] + dir[i] * *t;
    double last = 1, mg = INFINITY;
    int ok = 1;
    for (int i = 0; i < n - 1; ++i) {
        double b = dotn(p, m->pn + (f * (n - 1) + i) * n, n) - m->pd[f * (n - 1) + i];
        if (b <= 0 || b >= 1) ok = 0;
        last -= b;
        if (b < mg) mg = b;
    }
    if (last < mg) mg = last;
    if (!(last > 0)) ok = 0;
    *margin = mg;
    return ok;
}

static void ray_dir_normalized(const float* d, int n, double* out)
{
    for (int i = 0; i < n; ++i) out[i] = d[i];
    normalize_n(out, n); /* Ray ctor normalises, src/com/ray.h:34-36 */
}

void orc_trace_brute(const orc_mesh* m, int64_t count, const float* org, const float* dir, double eps_bary,
                     double rel_t, int32_t* idx, double* t, uint8_t* near)
{
    int n = m->n;
    for (int64_t r = 0; r < count; ++r) {
        double o[NM], d[NM];
        for (int i = 0; i < n; ++i) o[i] = org[r * n + i];
        ray_dir_normalized(dir + r * n, n, d);
        double best = INFINITY, second = INFINITY, amb = INFINITY;
        int64_t bi = -1;
        for (int64_t f = 0; f < m->F; ++f) {
            double tt, mg;
            int hit = simplex_hit(m, f, o, d, &tt, &mg);
            if (hit) {
                /* ray_intersection over a list: strict <, first index wins, ray_intersection.h:141-150 */
                if (tt < best) {
                    second = best;
                    best = tt;
                    bi = f;
                } else if (tt < second) {
                    second = tt;
                }
            }
            if (isfinite(mg) && fabs(mg) < eps_bary && tt > 0 && tt < amb) amb = tt;
        }
        idx[r] = (int32_t)bi;
        t[r] = bi >= 0 ? best : 0;
        int nr = 0;
        if (bi >= 0) {
            if (amb <= best * (1 + rel_t)) nr = 1;
            if (second - best <= rel_t * best) nr = 1;
        } else if (amb < INFINITY) {
            nr = 1;
        }
        near[r] = (uint8_t)nr;
    }
}

/* ------------------------------------------------------------------------ */
/* 2^N-tree                                                                   */

/* ParallelotopeOrtho planes for a node box: d1 = org + size, d2 = -org
 * (src/painter/space/parallelotope_ortho.h:207-229). */
static void node_box(const orc_mesh* m, int level, const int64_t* q, double* org, double* size)
{
    for (int i = 0; i < m->n; ++i) {
        double cell = ldexp(m->root_size[i], -level);
        org[i] = m->root_lo[i] + (double)q[i] * cell;
        size[i] = cell;
    }
}

/* Conservative box/simplex overlap: the separating-plane tests of
 * shapes_not_intersect_by_planes (src/painter/space/shape_intersection.h:150-189),
 * evaluated on the box centre and half sizes, with a guard delta. */
static int box_simplex_overlap(const orc_mesh* m, int64_t f, const double* org, const double* size)
{
    int n = m->n;
    double delta = m->dist_from_facet;
    double c[NM], h[NM];
    for (int i = 0; i < n; ++i) {
        h[i] = size[i] * 0.5;
        c[i] = org[i] + h[i];
    }
    const double* smn = m->smin + f * n;
    const double* smx = m->smax + f * n;
    for (int i = 0; i < n; ++i) {
        if (smn[i] > org[i] + size[i] + delta) return 0;
        if (smx[i] < org[i] - delta) return 0;
    }
    for (int k = 0; k < n; ++k) {
        const double* a = m->cons + (f * n + k) * (n + 1);
        double reach = fabs(a[0]) * h[0];
        for (int i = 1; i < n; ++i) reach = fma(fabs(a[i]), h[i], reach);
        double v = dotn(a, c, n) + a[n];
        if (v + reach < -delta) return 0;
    }
    {
        const double* a = m->ceq + f * (n + 1);
        double reach = fabs(a[0]) * h[0];
        for (int i = 1; i < n; ++i) reach = fma(fabs(a[i]), h[i], reach);
        double v = dotn(a, c, n) + a[n];
        if (fabs(v) > reach + delta) return 0;
    }
    return 1;
}

/* Node and entry arrays of a tree, or of one subtree built by a worker thread. */
typedef struct {
    onode* nodes;
    int64_t n_nodes, cap_nodes;
    int32_t* entries;
    int64_t n_entries, cap_entries;
} tbuf;

static int64_t push_node(tbuf* B, int n, int level, const int64_t* q)
{
    if (B->n_nodes == B->cap_nodes) {
        B->cap_nodes = B->cap_nodes ? B->cap_nodes * 2 : 1024;
        B->nodes = realloc(B->nodes, sizeof(onode) * B->cap_nodes);
    }
    onode* nd = &B->nodes[B->n_nodes];
    nd->level = level;
    memcpy(nd->q, q, sizeof(int64_t) * n);
    nd->child0 = -1;
    nd->begin = 0;
    nd->count = 0;
    return B->n_nodes++;
}

static void push_entries(tbuf* B, const int32_t* list, int64_t count, int64_t* begin)
{
    if (B->n_entries + count > B->cap_entries) {
        while (B->n_entries + count > B->cap_entries) B->cap_entries = B->cap_entries ? B->cap_entries * 2 : 4096;
        B->entries = realloc(B->entries, sizeof(int32_t) * B->cap_entries);
    }
    *begin = B->n_entries;
    memcpy(B->entries + B->n_entries, list, sizeof(int32_t) * count);
    B->n_entries += count;
}

/* The simplices of `list` that overlap child c of a node at (level, pq). */
static int64_t child_list(const orc_mesh* m, int level, const int64_t* pq, int c, const int32_t* list,
                          int64_t count, int32_t* out)
{
    int n = m->n;
    int64_t q[NM];
    for (int i = 0; i < n; ++i) q[i] = pq[i] * 2 + ((c >> i) & 1);
    double org[NM], size[NM];
    node_box(m, level + 1, q, org, size);
    int64_t k = 0;
    for (int64_t j = 0; j < count; ++j)
        if (box_simplex_overlap(m, list[j], org, size)) out[k++] = list[j];
    return k;
}

/* extend(), src/painter/space/tree.h:327-387, depth-first. */
static void build_rec(const orc_mesh* m, tbuf* B, int64_t node, int depth, const int32_t* list, int64_t count,
                      int min_objects)
{
    int n = m->n;
    if (depth >= m->max_depth || count <= min_objects) {
        int64_t b;
        push_entries(B, list, count, &b);
        B->nodes[node].begin = b;
        B->nodes[node].count = count;
        return;
    }
    int level = B->nodes[node].level;
    int64_t pq[NM];
    memcpy(pq, B->nodes[node].q, sizeof(int64_t) * n);
    int nch = 1 << n;
    int64_t first = -1;
    for (int c = 0; c < nch; ++c) {
        int64_t q[NM];
        for (int i = 0; i < n; ++i) q[i] = pq[i] * 2 + ((c >> i) & 1);
        int64_t id = push_node(B, n, level + 1, q);
        if (c == 0) first = id;
    }
    B->nodes[node].child0 = (int)first;
    int32_t* sub = malloc(sizeof(int32_t) * (count ? count : 1));
    for (int c = 0; c < nch; ++c) {
        int64_t k = child_list(m, level, pq, c, list, count, sub);
        int32_t* copy = malloc(sizeof(int32_t) * (k ? k : 1));
        memcpy(copy, sub, sizeof(int32_t) * k);
        build_rec(m, B, first + c, depth + 1, copy, k, min_objects);
        free(copy);
    }
    free(sub);
}

/* The reference builds the tree with worker threads (BoxJobs, tree.h:229-325,
 * 466-525).  Here the top levels are split on the calling thread (the
 * children of one split filtered in parallel), and the subtrees below them are
 * built by worker threads into buffers of their own and appended: the same
 * boxes and leaf lists as a depth-first build, in another node order. */
typedef struct {
    int64_t node;  /* in the main buffer */
    int level;
    int64_t q[NM];
    int depth;
    int32_t* list;
    int64_t count;
    tbuf sub;      /* the subtree below node (its local node 0 is node) */
} btask;

typedef struct {
    const orc_mesh* m;
    btask* tasks;
    int64_t ntasks;
    int64_t* next;
    int min_objects;
    /* child filtering of one split */
    int level;
    const int64_t* pq;
    const int32_t* list;
    int64_t count;
    int32_t** child;
    int64_t* child_count;
    int nch;
} bjob;

static void* split_worker(void* arg)
{
    bjob* j = arg;
    for (;;) {
        int64_t c = __atomic_fetch_add(j->next, 1, __ATOMIC_RELAXED);
        if (c >= j->nch) break;
        j->child[c] = malloc(sizeof(int32_t) * (j->count ? j->count : 1));
        j->child_count[c] = child_list(j->m, j->level, j->pq, (int)c, j->list, j->count, j->child[c]);
    }
    return NULL;
}

static void* subtree_worker(void* arg)
{
    bjob* j = arg;
    for (;;) {
        int64_t t = __atomic_fetch_add(j->next, 1, __ATOMIC_RELAXED);
        if (t >= j->ntasks) break;
        btask* k = &j->tasks[t];
        memset(&k->sub, 0, sizeof k->sub);
        push_node(&k->sub, j->m->n, k->level, k->q);
        build_rec(j->m, &k->sub, 0, k->depth, k->list, k->count, j->min_objects);
    }
    return NULL;
}

static void run_threads(int threads, void* (*fn)(void*), void* arg)
{
    pthread_t* th = calloc(threads, sizeof(pthread_t));
    for (int i = 0; i < threads; ++i) pthread_create(&th[i], NULL, fn, arg);
    for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    free(th);
}

static void build_parallel(orc_mesh* m, int min_objects, int threads)
{
    const int n = m->n, nch = 1 << n;
    tbuf M;
    memset(&M, 0, sizeof M);
    int64_t q0[NM] = {0};
    push_node(&M, n, 0, q0);
    btask* front = calloc(1, sizeof(btask));
    int64_t nfront = 1;
    front[0].node = 0;
    front[0].level = 0;
    front[0].depth = 1;
    front[0].count = m->F;
    front[0].list = malloc(sizeof(int32_t) * (m->F ? m->F : 1));
    for (int64_t f = 0; f < m->F; ++f) front[0].list[f] = (int32_t)f;
    /* split the top levels here until there are enough subtrees for the workers */
    for (int round = 0; round < 3 && nfront < 8 * threads; ++round) {
        btask* next = calloc(nfront * nch, sizeof(btask));
        int64_t nnext = 0;
        for (int64_t t = 0; t < nfront; ++t) {
            btask* k = &front[t];
            if (k->depth >= m->max_depth || k->count <= min_objects) {
                int64_t b;
                push_entries(&M, k->list, k->count, &b);
                M.nodes[k->node].begin = b;
                M.nodes[k->node].count = k->count;
                free(k->list);
                continue;
            }
            int64_t first = -1;
            for (int c = 0; c < nch; ++c) {
                int64_t q[NM];
                for (int i = 0; i < n; ++i) q[i] = k->q[i] * 2 + ((c >> i) & 1);
                int64_t id = push_node(&M, n, k->level + 1, q);
                if (c == 0) first = id;
            }
            M.nodes[k->node].child0 = (int)first;
            int32_t* child[1 << NM];
            int64_t cc[1 << NM];
            int64_t nx = 0;
            bjob j = {m, NULL, 0, &nx, min_objects, k->level, k->q, k->list, k->count, child, cc, nch};
            run_threads(threads < nch ? threads : nch, split_worker, &j);
            for (int c = 0; c < nch; ++c) {
                btask* d = &next[nnext++];
                d->node = first + c;
                d->level = k->level + 1;
                memcpy(d->q, M.nodes[first + c].q, sizeof(int64_t) * n);
                d->depth = k->depth + 1;
                d->list = child[c];
                d->count = cc[c];
            }
            free(k->list);
        }
        free(front);
        front = next;
        nfront = nnext;
    }
    /* the subtrees, on the workers */
    int64_t nx = 0;
    bjob j;
    memset(&j, 0, sizeof j);
    j.m = m;
    j.tasks = front;
    j.ntasks = nfront;
    j.next = &nx;
    j.min_objects = min_objects;
    run_threads(threads, subtree_worker, &j);
    /* append them in task order: node 0 of a subtree is its task's node */
    for (int64_t t = 0; t < nfront; ++t) {
        btask* k = &front[t];
        const int64_t off_n = M.n_nodes - 1, off_e = M.n_entries;
        onode* root = &M.nodes[k->node];
        const onode* r = &k->sub.nodes[0];
        root->child0 = r->child0 >= 0 ? (int)(r->child0 + off_n) : -1;
        root->begin = r->begin + off_e;
        root->count = r->count;
        for (int64_t i = 1; i < k->sub.n_nodes; ++i) {
            onode nd = k->sub.nodes[i];
            if (nd.child0 >= 0) nd.child0 = (int)(nd.child0 + off_n);
            else nd.begin += off_e;
            int64_t id = push_node(&M, n, nd.level, nd.q);
            M.nodes[id] = nd;
        }
        int64_t b;
        push_entries(&M, k->sub.entries, k->sub.n_entries, &b);
        free(k->sub.nodes);
        free(k->sub.entries);
        free(k->list);
    }
    free(front);
    m->nodes = M.nodes;
    m->n_nodes = M.n_nodes;
    m->cap_nodes = M.cap_nodes;
    m->entries = M.entries;
    m->n_entries = M.n_entries;
    m->cap_entries = M.cap_entries;
}

int orc_box_simplex_overlap(const orc_mesh* m, int64_t f, const double* org, const double* size)
{
    if (!m->has_tree || f < 0 || f >= m->F) return -1;
    return box_simplex_overlap(m, f, org, size);
}

void orc_tree_root(const orc_mesh* m, double* lo, double* size)
{
    memcpy(lo, m->root_lo, sizeof(double) * m->n);
    memcpy(size, m->root_size, sizeof(double) * m->n);
}

int orc_tree_build(orc_mesh* m, int max_depth, int min_objects)
{
    return orc_tree_build_threads(m, max_depth, min_objects, 1);
}

int orc_tree_build_threads(orc_mesh* m, int max_depth, int min_objects, int threads)
{
    if (threads < 1) return fail("tree build thread count (%d) must be positive", threads);
    int n = m->n;
    /* decompose(), tree.h:466-525 and min_max_and_distance(), tree.h:167-226 */
    if (!(max_depth >= 1 && max_depth <= 10) || !(min_objects >= 2 && min_objects <= 100))
        return fail("Error limits for spatial subdivision tree");
    const double eps = 2.220446049250313e-16;
    double mn[NM], mx[NM], all_max = -INFINITY;
    for (int i = 0; i < n; ++i) {
        mn[i] = m->mn[i];
        mx[i] = m->mx[i];
        if (!(mn[i] < mx[i])) return fail("Objects for (2^N)-tree don't form N-dimensional object");
    }
    for (int i = 0; i < n; ++i) {
        double abs_max = fmax(fabs(mn[i]), fabs(mx[i]));
        double guard = abs_max * (20 * eps);
        mn[i] -= guard;
        mx[i] += guard;
        all_max = fmax(abs_max, all_max);
    }
    double dist = all_max * (20 * eps);
    int max_div = 1 << (max_depth - 1);
    for (int i = 0; i < n; ++i)
        if (dist >= (mx[i] - mn[i]) / max_div / 2)
            return fail("The minimal distance from facets is greater than one half of the minimum box size");
    for (int i = 0; i < n; ++i) {
        m->root_lo[i] = mn[i];
        m->root_size[i] = mx[i] - mn[i];
    }
    m->dist_from_facet = dist;
    m->max_depth = max_depth;
    /* per-simplex overlap data: wrapper vertices min/max and constraints,
     * src/painter/space/hyperplane_simplex_wrapper.h:58-65 */
    free(m->cons); free(m->ceq); free(m->smin); free(m->smax);
    m->cons = malloc(sizeof(double) * m->F * n * (n + 1));
    m->ceq = malloc(sizeof(double) * m->F * (n + 1));
    m->smin = malloc(sizeof(double) * m->F * n);
    m->smax = malloc(sizeof(double) * m->F * n);
    for (int64_t f = 0; f < m->F; ++f) {
        double verts[NM * NM];
        for (int k = 0; k < n; ++k) memcpy(verts + k * n, m->vtx + (int64_t)m->fac[f * n + k] * n, n * sizeof(double));
        simplex_constraints(n, m->normal + f * n, verts, m->pn + f * (n - 1) * n, m->pd + f * (n - 1),
                            m->cons + f * n * (n + 1), m->ceq + f * (n + 1));
        for (int i = 0; i < n; ++i) {
            double a = INFINITY, b = -INFINITY;
            for (int k = 0; k < n; ++k) {
                a = fmin(a, verts[k * n + i]);
                b = fmax(b, verts[k * n + i]);
            }
            m->smin[f * n + i] = a;
            m->smax[f * n + i] = b;
        }
    }
    free(m->nodes); m->nodes = NULL; m->n_nodes = m->cap_nodes = 0;
    free(m->entries); m->entries = NULL; m->n_entries = m->cap_entries = 0;
    build_parallel(m, min_objects, threads);
    m->has_tree = 1;
    return 0;
}

void orc_tree_info(const orc_mesh* m, int64_t* nodes, int64_t* leaves, int64_t* entries, int* depth)
{
    int64_t lv = 0;
    int d = 0;
    for (int64_t i = 0; i < m->n_nodes; ++i) {
        if (m->nodes[i].child0 < 0) ++lv;
        if (m->nodes[i].level > d) d = m->nodes[i].level;
    }
    *nodes = m->n_nodes;
    *leaves = lv;
    *entries = m->n_entries;
    *depth = d;
}

int64_t orc_tree_dump(const orc_mesh* m, int64_t* out, int64_t cap)
{
    int n = m->n;
    int64_t w = 0;
    for (int64_t i = 0; i < m->n_nodes; ++i) {
        const onode* nd = &m->nodes[i];
        if (nd->child0 >= 0) continue;
        if (out && w + 2 + n + nd->count <= cap) {
            out[w] = nd->level;
            for (int k = 0; k < n; ++k) out[w + 1 + k] = nd->q[k];
            out[w + 1 + n] = nd->count;
            for (int64_t j = 0; j < nd->count; ++j) out[w + 2 + n + j] = m->entries[nd->begin + j];
        }
        w += 2 + n + nd->count;
    }
    return w;
}

/* ParallelotopeOrtho::intersect_impl, parallelotope_ortho.h:254-313 */
static int box_intersect_impl(int n, const double* d1, const double* d2, const double* org, const double* dir,
                              double* first, double* second)
{
    double f_max = -INFINITY, b_min = INFINITY;
    for (int i = 0; i < n; ++i) {
        double s = dir[i];
        if (s == 0) {
            double d = org[i];
            if (d - d1[i] > 0 || -d - d2[i] > 0) return 0;
            continue;
        }
        double d = org[i];
        double a1 = (d1[i] - d) / s;
        double a2 = (d2[i] + d) / -s;
        if (s < 0) {
            f_max = fmax(a1, f_max);
            b_min = fmin(a2, b_min);
        } else {
            b_min = fmin(a1, b_min);
            f_max = fmax(a2, f_max);
        }
        if (b_min <= 0 || b_min < f_max) return 0;
    }
    *first = f_max;
    *second = b_min;
    return 1;
}

static void node_planes(const orc_mesh* m, const onode* nd, double* d1, double* d2)
{
    double org[NM], size[NM];
    node_box(m, nd->level, nd->q, org, size);
    for (int i = 0; i < m->n; ++i) {
        d1[i] = org[i] + size[i];
        d2[i] = -org[i];
    }
}

/* ParallelotopeOrtho::inside, parallelotope_ortho.h:366-380 */
static int box_inside(int n, const double* d1, const double* d2, const double* p)
{
    for (int i = 0; i < n; ++i)
        if (!(p[i] <= d1[i]) || !(-p[i] <= d2[i])) return 0;
    return 1;
}

/* ParallelotopeOrtho::normal, parallelotope_ortho.h:340-364 */
static void box_normal(int n, const double* d1, const double* d2, const double* p, double* nrm)
{
    double mn = INFINITY;
    for (int i = 0; i < n; ++i) {
        double l = fabs(p[i] - d1[i]);
        if (l < mn) {
            mn = l;
            memset(nrm, 0, sizeof(double) * n);
            nrm[i] = 1;
        }
        l = fabs(p[i] + d2[i]);
        if (l < mn) {
            mn = l;
            memset(nrm, 0, sizeof(double) * n);
            nrm[i] = -1;
        }
    }
}

/* find_box_for_point, tree.h:441-463 */
static int64_t find_box(const orc_mesh* m, int64_t node, const double* p)
{
    int n = m->n;
    double d1[NM], d2[NM];
    node_planes(m, &m->nodes[node], d1, d2);
    if (!box_inside(n, d1, d2, p)) return -1;
    if (m->nodes[node].child0 < 0) return node;
    for (int c = 0; c < (1 << n); ++c) {
        int64_t r = find_box(m, m->nodes[node].child0 + c, p);
        if (r >= 0) return r;
    }
    return -1;
}

/* Mesh::intersect_approximate + intersect_precise, src/painter/shapes/mesh.cpp:137-166,
 * SpatialSubdivisionTree::trace_ray, tree.h:534-590. */
static int trace_tree_one(const orc_mesh* m, const double* org0, const double* dir, double* t_out, int64_t* f_out)
{
    int n = m->n;
    double d1[NM], d2[NM];
    node_planes(m, &m->nodes[0], d1, d2);
    double first, second;
    if (!box_intersect_impl(n, d1, d2, org0, dir, &first, &second)) return 0;
    double root_t = first > 0 ? first : second;
    double org[NM], interior[NM];
    memcpy(org, org0, sizeof(double) * n);
    memcpy(interior, org0, sizeof(double) * n);
    int is_first = 1;
    for (int guard = 0; guard < 1000000; ++guard) {
        int64_t box = find_box(m, 0, interior);
        if (box >= 0) {
            const onode* nd = &m->nodes[box];
            node_planes(m, nd, d1, d2);
            if (nd->count > 0) {
                /* ray_intersection over the leaf list with the original ray, ray_intersection.h:134-159 */
                double best = INFINITY;
                int64_t bf = -1;
                for (int64_t j = 0; j < nd->count; ++j) {
                    int64_t f = m->entries[nd->begin + j];
                    double tt, mg;
                    if (simplex_hit(m, f, org0, dir, &tt, &mg) && tt < best) {
                        best = tt;
                        bf = f;
                    }
                }
                if (bf >= 0) {
                    double p[NM];
                    for (int i = 0; i < n; ++i) p[i] = org0[i] + dir[i] * best;
                    if (box_inside(n, d1, d2, p)) {
                        *t_out = best;
                        *f_out = bf;
                        return 1;
                    }
                }
            }
            double t;
            if (!box_intersect_impl(n, d1, d2, org, dir, &first, &second)) return 0;
            t = second;
            double ip[NM], nrm[NM];
            for (int i = 0; i < n; ++i) ip[i] = org[i] + dir[i] * t;
            memcpy(org, ip, sizeof(double) * n);
            box_normal(n, d1, d2, ip, nrm);
            for (int i = 0; i < n; ++i) interior[i] = ip[i] + m->dist_from_facet * nrm[i];
        } else {
            if (!is_first) return 0;
            double ip[NM], nrm[NM];
            for (int i = 0; i < n; ++i) ip[i] = org[i] + dir[i] * root_t;
            memcpy(org, ip, sizeof(double) * n);
            node_planes(m, &m->nodes[0], d1, d2);
            box_normal(n, d1, d2, ip, nrm);
            for (int i = 0; i < n; ++i) interior[i] = ip[i] - m->dist_from_facet * nrm[i];
        }
        is_first = 0;
    }
    return 0;
}

void orc_trace_tree(const orc_mesh* m, int64_t count, const float* org, const float* dir, int32_t* idx, double* t)
{
    int n = m->n;
    for (int64_t r = 0; r < count; ++r) {
        double o[NM], d[NM];
        for (int i = 0; i < n; ++i) o[i] = org[r * n + i];
        ray_dir_normalized(dir + r * n, n, d);
        double tt;
        int64_t f;
        if (m->has_tree && trace_tree_one(m, o, d, &tt, &f)) {
            idx[r] = (int32_t)f;
            t[r] = tt;
        } else {
            idx[r] = -1;
            t[r] = 0;
        }
    }
}

/* Nearest hit used by the renderer: the tree when built, else brute force. */
static int nearest(const orc_mesh* m, const double* org, const double* dir, double* t, int64_t* f)
{
    if (m->has_tree) return trace_tree_one(m, org, dir, t, f);
    double best = INFINITY;
    int64_t bf = -1;
    for (int64_t k = 0; k < m->F; ++k) {
        double tt, mg;
        if (simplex_hit(m, k, org, dir, &tt, &mg) && tt < best) {
            best = tt;
            bf = k;
        }
    }
    if (bf < 0) return 0;
    *t = best;
    *f = bf;
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Painter                                                                    */

int orc_samples_per_dimension(int dims, int spp)
{
    /* StratifiedJitteredSampler::one_dimension_size, sampler.h:31-52 */
    if (spp < 1) return -1;
    double v = pow(spp, 1.0 / dims);
    unsigned f = (unsigned)floor(v), c = (unsigned)ceil(v);
    unsigned long long p = 1;
    for (int i = 0; i < dims; ++i) p *= f;
    if (p >= (unsigned)spp) return (int)f;
    p = 1;
    for (int i = 0; i < dims; ++i) p *= c;
    if (p >= (unsigned)spp) return (int)c;
    return -1;
}

void orc_cosine_hemisphere(int n, const double* normal, const double* u, double* res)
{
    /* random_cosine_weighted_on_hemisphere, src/painter/sampling/sphere.h:87-116:
     * tangent part uniform in direction with |v|^2 ~ U(0,1) after the N>=4
     * reweighting, normal part sqrt(1-|v|^2). */
    const double TWO_PI = 6.283185307179586476925286766559;
    double dirv[NM - 1];
    int m = n - 1;
    if (m == 2) {
        double phi = TWO_PI * u[1];
        dirv[0] = cos(phi);
        dirv[1] = sin(phi);
    } else if (m == 3) {
        double z = 1 - 2 * u[1];
        double rho = sqrt(fmax(0.0, 1 - z * z));
        double phi = TWO_PI * u[2];
        dirv[0] = rho * cos(phi);
        dirv[1] = rho * sin(phi);
        dirv[2] = z;
    } else if (m == 4) {
        double a = u[1];
        double p1 = TWO_PI * u[2], p2 = TWO_PI * u[3];
        double ra = sqrt(a), rb = sqrt(1 - a);
        dirv[0] = ra * cos(p1);
        dirv[1] = ra * sin(p1);
        dirv[2] = rb * cos(p2);
        dirv[3] = rb * sin(p2);
    } else {
        double g[6];
        for (int k = 0; k < 3; ++k) {
            double r = sqrt(-2 * log(1 - u[1 + 2 * k]));
            double th = TWO_PI * u[2 + 2 * k];
            g[2 * k] = r * cos(th);
            g[2 * k + 1] = r * sin(th);
        }
        double l2 = 0;
        for (int k = 0; k < m; ++k) l2 += g[k] * g[k];
        double l = sqrt(l2);
        for (int k = 0; k < m; ++k) dirv[k] = g[k] / l;
    }
    double tl = sqrt(u[0]);
    double nc = sqrt(1 - u[0]);
    double basis[(NM - 1) * NM];
    orc_orthogonal_complement(n, normal, basis);
    for (int i = 0; i < n; ++i) res[i] = nc * normal[i];
    for (int k = 0; k < m; ++k) {
        double v = tl * dirv[k];
        for (int i = 0; i < n; ++i) res[i] = res[i] + v * basis[k * n + i];
    }
}

/* A shape's planes (FP64): Parallelotope::create_planes (parallelotope.h:129-153);
 * HyperplaneParallelotope::set_data and HyperplaneSimplexGeometry::set_data
 * (hyperplane_parallelotope.h:68-74, hyperplane_geometry.h:80-110). */
typedef struct {
    int kind;
    double org[NM];
    double n[NM];
    double pn[NM][NM];
    double d1[NM], d2[NM];
} shape_geom;

static void shape_setup(int n, int kind, const double* org, const double (*edges)[NM], shape_geom* g)
{
    g->kind = kind;
    memcpy(g->org, org, sizeof(double) * n);
    double vec[(NM - 1) * NM];
    if (kind == 0) {
        for (int i = 0; i < n; ++i) {
            for (int a = 0, row = 0; a < n; ++a)
                if (a != i) {
                    memcpy(vec + row * n, edges[a], sizeof(double) * n);
                    ++row;
                }
            double nn[NM], ov[NM];
            orc_ortho_nn(n, vec, nn);
            normalize_n(nn, n);
            for (int k = 0; k < n; ++k) ov[k] = org[k] + edges[i][k];
            double d1 = dotn(org, nn, n), d2 = -dotn(ov, nn, n);
            if (dotn(nn, edges[i], n) > 0) {
                for (int k = 0; k < n; ++k) nn[k] = -nn[k];
                d1 = -d1;
                d2 = -d2;
            }
            memcpy(g->pn[i], nn, sizeof(double) * n);
            g->d1[i] = d1;
            g->d2[i] = d2;
        }
        return;
    }
    for (int a = 0; a < n - 1; ++a) memcpy(vec + a * n, edges[a], sizeof(double) * n);
    orc_ortho_nn(n, vec, g->n);
    normalize_n(g->n, n);
    double pts[NM][NM];
    for (int a = 0; a < n - 1; ++a)
        for (int k = 0; k < n; ++k) pts[a][k] = org[k] + edges[a][k];
    memcpy(pts[n - 1], org, sizeof(double) * n);
    for (int a = 0; a < n - 1; ++a)
        for (int k = 0; k < n; ++k) vec[a * n + k] = pts[a][k] - pts[n - 1][k];
    for (int i = 0; i < n - 1; ++i) {
        double m[(NM - 1) * NM], pn[NM];
        memcpy(m, vec, sizeof(double) * (n - 1) * n);
        memcpy(m + i * n, g->n, sizeof(double) * n);
        orc_ortho_nn(n, m, pn);
        double d = dotn(pts[n - 1], pn, n);
        double dist = dotn(pts[i], pn, n) - d;
        for (int k = 0; k < n; ++k) g->pn[i][k] = pn[k] / dist;
        g->d1[i] = d / dist;
    }
}

/* Parallelotope::intersect (parallelotope.h:196-246); HyperplaneParallelotopeGeometry::
 * intersect (hyperplane_geometry.h:41-52, 240-261) */
static int shape_intersect(int n, const shape_geom* g, const double* o, const double* d, double* t)
{
    if (g->kind == 0) {
        double f_max = -DBL_MAX, b_min = DBL_MAX;
        for (int i = 0; i < n; ++i) {
            double sd = dotn(d, g->pn[i], n);
            double od = dotn(o, g->pn[i], n);
            if (sd == 0) {
                if (od - g->d1[i] > 0 || -od - g->d2[i] > 0) return 0;
                continue;
            }
            double a1 = (g->d1[i] - od) / sd;
            double a2 = (g->d2[i] + od) / -sd;
            if (sd < 0) {
                if (a1 > f_max) f_max = a1;
                if (a2 < b_min) b_min = a2;
            } else {
                if (a1 < b_min) b_min = a1;
                if (a2 > f_max) f_max = a2;
            }
            if (b_min <= 0 || b_min < f_max) return 0;
        }
        *t = f_max > 0 ? f_max : b_min;
        return 1;
    }
    double sd = dotn(g->n, d, n);
    if (sd == 0) return 0;
    double v[NM], p[NM];
    for (int k = 0; k < n; ++k) v[k] = g->org[k] - o[k];
    double tt = dotn(v, g->n, n) / sd;
    if (!(tt > 0)) return 0;
    for (int k = 0; k < n; ++k) p[k] = o[k] + d[k] * tt;
    for (int i = 0; i < n - 1; ++i) {
        double c = dotn(p, g->pn[i], n) - g->d1[i];
        if (c <= 0 || c >= 1) return 0;
    }
    *t = tt;
    return 1;
}

/* Parallelotope::normal (parallelotope.h:261-290), HyperplaneParallelotope::normal */
static void shape_normal(int n, const shape_geom* g, const double* p, double* out)
{
    if (g->kind != 0) {
        memcpy(out, g->n, sizeof(double) * n);
        return;
    }
    double mn = DBL_MAX;
    int best = 0, neg = 0;
    for (int i = 0; i < n; ++i) {
        double dd = dotn(p, g->pn[i], n);
        double l = fabs(dd - g->d1[i]);
        if (l < mn) {
            mn = l;
            best = i;
            neg = 0;
        }
        l = fabs(-dd - g->d2[i]);
        if (l < mn) {
            mn = l;
            best = i;
            neg = 1;
        }
    }
    for (int k = 0; k < n; ++k) out[k] = neg ? -g->pn[best][k] : g->pn[best][k];
}

int orc_shape_intersect(int n, int kind, const double* org, const double* edges, const double* o, const double* d,
                        double* t, double* normal)
{
    shape_geom g;
    shape_setup(n, kind, org, (const double(*)[NM])edges, &g);
    if (!shape_intersect(n, &g, o, d, t)) return 0;
    double p[NM];
    for (int k = 0; k < n; ++k) p[k] = o[k] + d[k] * *t;
    shape_normal(n, &g, p, normal);
    return 1;
}

typedef struct {
    const orc_mesh* m;
    const orc_scene* s;
    shape_geom shapes[ORC_MAX_SHAPES];
    double cam_dir[NM];        /* perspective: times the screen distance */
    double axes[NM - 1][NM];   /* parallel: times units_per_pixel */
    double square_radius;      /* spherical */
    double ray_offset;
    double coef;
    double light_coef;
    float bg_light[3];
    int k;  /* samples per dimension */
    int nsamples;
} pctx;

typedef struct {
    int64_t pixel;
    int sample;
    int64_t rays;
} pstate;

static float maxf3(const float* c)
{
    float r = c[0];
    if (c[1] > r) r = c[1];
    if (c[2] > r) r = c[2];
    return r;
}

/* SurfaceProperties for a mesh hit: geometric and shading normal,
 * src/painter/visible_shapes.h:148-162, mesh_hyperplane_simplex.cpp:151-186. */
static void hit_normals(const orc_mesh* m, int64_t f, const double* p, double* gn, double* sn)
{
    int n = m->n;
    memcpy(gn, m->normal + f * n, sizeof(double) * n);
    normalize_n(gn, n);  /* set_geometric_normal normalises, objects.h:72-75 */
    if (!m->vn || m->ntype[f] == 0) {
        memcpy(sn, gn, sizeof(double) * n);
        return;
    }
    /* interpolate(), hyperplane_geometry.h:206-219 */
    double bc[NM];
    bc[n - 1] = 1;
    for (int i = 0; i < n - 1; ++i) {
        bc[i] = dotn(p, m->pn + (f * (n - 1) + i) * n, n) - m->pd[f * (n - 1) + i];
        bc[n - 1] -= bc[i];
    }
    double r[NM];
    for (int k = 0; k < n; ++k) {
        const double* vn = m->vn + (int64_t)m->fnrm[f * n + k] * n;
        double sgn = (m->ntype[f] == 2 && m->rev[f * n + k]) ? -1.0 : 1.0;
        for (int i = 0; i < n; ++i) {
            double term = bc[k] * (sgn * vn[i]);
            r[i] = (k == 0) ? term : r[i] + term;
        }
    }
    normalize_n(r, n);
    memcpy(sn, r, sizeof(double) * n);
    normalize_n(sn, n);  /* set_shading_normal normalises */
}

static int light_visible(const pctx* c, pstate* st, const double* org, const double* dir, double dist)
{
    /* light_source_is_visible / object_is_obstacle_to_light, painter.cpp:142-186:
     * the objects in order (the mesh, then the shapes), one ray count each,
     * until one blocks the light */
    st->rays += 1;
    double t;
    int64_t f;
    if (nearest(c->m, org, dir, &t, &f) && t < dist) return 0;
    for (int k = 0; k < c->s->shape_count; ++k) {
        st->rays += 1;
        if (shape_intersect(c->m->n, &c->shapes[k], org, dir, &t) && t < dist) return 0;
    }
    return 1;
}

/* ray_intersection over the objects (src/painter/space/ray_intersection.h:20-80):
 * the nearest of the mesh hit and the shapes' hits; *shape = -1 for the mesh */
static int nearest_object(const pctx* c, const double* org, const double* dir, double* t, int64_t* f, int* shape)
{
    int found = nearest(c->m, org, dir, t, f);
    *shape = -1;
    double best = found ? *t : DBL_MAX;
    for (int k = 0; k < c->s->shape_count; ++k) {
        double ts;
        if (shape_intersect(c->m->n, &c->shapes[k], org, dir, &ts) && ts < best) {
            best = ts;
            *shape = k;
            found = 1;
        }
    }
    if (found) *t = best;
    return found;
}

static void trace_path(const pctx* c, pstate* st, int level, float color_level, const double* org,
                       const double* dir, int diffuse_reflection, float* out);

/* One light of direct_diffuse_lighting (painter.cpp:205-289) with
 * PointLight::properties or ConstantLight::properties
 * (src/painter/lights/light_source.h:26-78); adds its term to out. */
static void direct_one(const pctx* c, pstate* st, const double* p, const double* gn, const double* sn, int mesh,
                       int kind, const double* pos, const float* rgb, double coef, float* out)
{
    int n = c->m->n;
    const orc_scene* s = c->s;
    double vl[NM];
    for (int i = 0; i < n; ++i) vl[i] = pos[i] - p[i];
    double sq = dotn(vl, vl, n);
    double scale = kind == 1 ? 1.0 : (n == 3) ? coef / sq : coef / pow(sq, (double)(n - 1) / 2);
    float lc[3];
    for (int k = 0; k < 3; ++k) lc[k] = rgb[k] * (float)scale;
    if (maxf3(lc) < 1e-4f) return;
    double ldir[NM];
    memcpy(ldir, vl, sizeof ldir);
    normalize_n(ldir, n);
    double dln = dotn(ldir, sn, n);
    if (dln <= 0) return;
    double dist = sqrt(dotn(vl, vl, n));
    double o[NM];
    int mesh_smooth_back = mesh && c->m->vn && s->smooth_normal && dotn(ldir, gn, n) < 0;
    if (!mesh_smooth_back) {
        for (int i = 0; i < n; ++i) o[i] = p[i] + ldir[i] * c->ray_offset;
        if (!light_visible(c, st, o, ldir, dist)) return;
    } else {
        st->rays += 1;
        for (int i = 0; i < n; ++i) o[i] = p[i] + ldir[i] * c->ray_offset;
        double t;
        int64_t f;
        if (!nearest(c->m, o, ldir, &t, &f)) return;
        if (t >= dist) return;
        st->rays += 1;
        double rd[NM], ro[NM];
        for (int i = 0; i < n; ++i) {
            rd[i] = -ldir[i];
            ro[i] = o[i] + rd[i] * (2 * c->ray_offset);
        }
        double tr;
        if (nearest(c->m, ro, rd, &tr, &f) && tr < t) return;
        double o2[NM];
        for (int i = 0; i < n; ++i) o2[i] = o[i] + ldir[i] * (t + c->ray_offset);
        if (!light_visible(c, st, o2, ldir, dist - t)) return;
    }
    float w = (float)(c->coef * dln);
    for (int k = 0; k < 3; ++k) out[k] = out[k] + lc[k] * w;
}

/* direct_diffuse_lighting, painter.cpp:197-292: the lights in order */
static void direct_lighting(const pctx* c, pstate* st, const double* p, const double* gn, const double* sn,
                            int mesh, float* out)
{
    const orc_scene* s = c->s;
    int n = c->m->n;
    out[0] = out[1] = out[2] = 0;
    if (!s->no_light)
        direct_one(c, st, p, gn, sn, mesh, s->light_kind, s->light_pos, s->light_rgb, c->light_coef, out);
    for (int l = 0; l < s->extra_light_count && l < ORC_MAX_LIGHTS - 1; ++l)
        direct_one(c, st, p, gn, sn, mesh, s->extra_lights[l].kind, s->extra_lights[l].pos, s->extra_lights[l].rgb,
                   pow(s->extra_lights[l].unit_distance, n - 1), out);
}

/* trace_path, painter.cpp:325-397 and diffuse_lighting, :298-323 */
static void trace_path(const pctx* c, pstate* st, int level, float color_level, const double* org,
                       const double* dir, int diffuse_reflection, float* out)
{
    const orc_mesh* m = c->m;
    const orc_scene* s = c->s;
    int n = m->n;
    st->rays += 1;
    out[0] = out[1] = out[2] = 0;
    double t;
    int64_t f;
    int shape;
    if (!nearest_object(c, org, dir, &t, &f, &shape)) {
        const float* bg = (diffuse_reflection && !s->background_not_light) ? c->bg_light : s->background_rgb;
        out[0] = bg[0]; out[1] = bg[1]; out[2] = bg[2];
        return;
    }
    const int mesh = shape < 0;
    double p[NM], gn[NM], sn[NM];
    for (int i = 0; i < n; ++i) p[i] = org[i] + dir[i] * t;
    const float* rgb = s->object_rgb;
    float dif = s->diffuse;
    if (mesh) {
        hit_normals(m, f, p, gn, sn);
    } else {
        shape_normal(n, &c->shapes[shape], p, gn);
        normalize_n(gn, n);  /* set_geometric_normal normalises, objects.h:72-75 */
        memcpy(sn, gn, sizeof(double) * n);
        rgb = s->shapes[shape].rgb;
        dif = s->shapes[shape].diffuse;
    }
    double dgn = dotn(dir, gn, n);
    if (fabs(dgn) <= 0) return;
    if (!mesh && s->shapes[shape].light) {
        /* trace_path, painter.cpp:355-358 */
        const float* lc = diffuse_reflection ? s->shapes[shape].light_rgb : s->shapes[shape].rgb;
        out[0] = lc[0]; out[1] = lc[1]; out[2] = lc[2];
        return;
    }
    if (!(mesh && m->vn && s->smooth_normal)) memcpy(sn, gn, sizeof(double) * n);
    if (dgn > 0) {
        for (int i = 0; i < n; ++i) { gn[i] = -gn[i]; sn[i] = -sn[i]; }
    }
    if (!(dif > 0)) return;
    float sc[3];
    for (int k = 0; k < 3; ++k) sc[k] = dif * rgb[k];
    float new_level = color_level * maxf3(sc);
    if (!(new_level >= 1e-4f)) return;
    float direct[3];
    direct_lighting(c, st, p, gn, sn, mesh, direct);
    float diffuse[3] = {0, 0, 0};
    if (level < s->max_depth) {
        double u[8];
        uniforms4(s->seed, st->pixel, st->sample, s->pass_index, level, 2, u);
        if (n >= 6) uniforms4(s->seed, st->pixel, st->sample, s->pass_index, level, 3, u + 4);
        double nd[NM];
        orc_cosine_hemisphere(n, sn, u, nd);
        normalize_n(nd, n);  /* Ray ctor */
        if (!mesh || dotn(nd, gn, n) > 0) {  /* only a mesh's bounce must leave its side, painter.cpp:308-313 */
            double o[NM];
            for (int i = 0; i < n; ++i) o[i] = p[i] + nd[i] * c->ray_offset;
            trace_path(c, st, level + 1, new_level, o, nd, 1, diffuse);
        }
    }
    for (int k = 0; k < 3; ++k) out[k] = out[k] + sc[k] * (direct[k] + diffuse[k]);
}

/* Projector constructors (src/painter/projectors/projector.h): unit_dir_and_axes
 * (:68-86), then PerspectiveProjector (:110-136): camera_dir times
 * screen_size[0]/2 * tan(pi/2 - half_angle); ParallelProjector (:161-185): axes
 * times units_per_pixel; SphericalProjector (:212-249): square_radius =
 * (screen_size[0]/2 / sin(half_angle))^2, with its view-angle checks. */
static int projector_setup(const orc_scene* s, int n, double* cam_dir, double axes[][NM], double* square_radius)
{
    memcpy(cam_dir, s->cam_dir, sizeof(double) * n);
    normalize_n(cam_dir, n);
    for (int a = 0; a < n - 1; ++a) {
        memcpy(axes[a], s->axes[a], sizeof(double) * n);
        normalize_n(axes[a], n);
    }
    const double PI_D = 3.14159265358979323846;
    const double half_angle = s->view_angle_degrees * 0.5 * PI_D / 180;
    *square_radius = 0;
    if (s->projector == 0) {
        if (!(s->units_per_pixel > 0)) return fail("Error units per pixel for parallel projection");
        for (int a = 0; a < n - 1; ++a)
            for (int i = 0; i < n; ++i) axes[a][i] = axes[a][i] * s->units_per_pixel;
    } else if (s->projector == 1) {
        if (!(s->view_angle_degrees > 0 && s->view_angle_degrees < 180))
            return fail("Perspective projection: error view angle %g", s->view_angle_degrees);
        double dir_length = s->screen[0] * 0.5 * tan(PI_D / 2 - half_angle);
        for (int i = 0; i < n; ++i) cam_dir[i] = cam_dir[i] * dir_length;
    } else if (s->projector == 2) {
        double sin_alpha = sin(half_angle);
        if (!(s->view_angle_degrees > 0)) return fail("Spherical projection view angle is not positive");
        double k = sin_alpha / s->screen[0];
        double r = sin_alpha * sin_alpha;
        for (int i = 1; i < n - 1; ++i) r += (k * s->screen[i]) * (k * s->screen[i]);
        if (!(r < 1)) return fail("Spherical projection view angle %g is too big", s->view_angle_degrees);
        double rad = s->screen[0] * 0.5 / sin_alpha;
        *square_radius = rad * rad;
    } else {
        return fail("unknown projector %d", s->projector);
    }
    return 0;
}

/* Projector::ray (projector.h:143-150, 194-201, 259-274) at screen point
 * sp = screen_org + point; the direction is normalised as Ray's constructor
 * does (src/com/ray.h:34-36). */
static void projector_ray(const orc_scene* s, int n, const double* cam_dir, double axes[][NM], double square_radius,
                          const double* sp, double* org, double* dir)
{
    double sd[NM];
    for (int j = 0; j < n; ++j) {
        sd[j] = axes[0][j] * sp[0];
        for (int i = 1; i < n - 1; ++i) sd[j] = sd[j] + axes[i][j] * sp[i];
    }
    if (s->projector == 0) {
        for (int j = 0; j < n; ++j) {
            org[j] = s->cam_org[j] + sd[j];
            dir[j] = cam_dir[j];
        }
        return;
    }
    double z = 1;
    if (s->projector == 2) z = sqrt(square_radius - dotn(sp, sp, n - 1));
    for (int j = 0; j < n; ++j) {
        org[j] = s->cam_org[j];
        dir[j] = cam_dir[j] * z + sd[j];
    }
    normalize_n(dir, n);
}

int orc_projector_ray(const orc_scene* s, const double* point, double* org, double* dir)
{
    int n = s->n;
    double cam_dir[NM], axes[NM - 1][NM], r2, sp[NM - 1];
    if (projector_setup(s, n, cam_dir, axes, &r2)) return 1;
    for (int i = 0; i < n - 1; ++i) sp[i] = -s->screen[i] * 0.5 + point[i];
    projector_ray(s, n, cam_dir, axes, r2, sp, org, dir);
    return 0;
}

static int make_ctx(const orc_mesh* m, const orc_scene* s, pctx* c)
{
    int n = m->n;
    if (s->n != n) return fail("scene dimension %d != mesh dimension %d", s->n, n);
    c->m = m;
    c->s = s;
    if (projector_setup(s, n, c->cam_dir, c->axes, &c->square_radius)) return 1;
    /* compute_ray_offset, painter.cpp:539-558, over every object's vertices */
    double all_max = -INFINITY;
    for (int i = 0; i < n; ++i) all_max = fmax(all_max, fmax(fabs(m->mn[i]), fabs(m->mx[i])));
    if (s->shape_count < 0 || s->shape_count > ORC_MAX_SHAPES) return fail("bad shape count %d", s->shape_count);
    for (int k = 0; k < s->shape_count; ++k) {
        shape_setup(n, s->shapes[k].kind, s->shapes[k].org, (const double(*)[NM])s->shapes[k].edges, &c->shapes[k]);
        int edges = s->shapes[k].kind == 0 ? n : n - 1;
        for (int mask = 0; mask < (1 << edges); ++mask)
            for (int i = 0; i < n; ++i) {
                double x = s->shapes[k].org[i];
                for (int a = 0; a < edges; ++a)
                    if (mask & (1 << a)) x += s->shapes[k].edges[a][i];
                all_max = fmax(all_max, fabs(x));
            }
    }
    c->ray_offset = all_max * (1000 * 2.220446049250313e-16);
    c->coef = orc_cosine_sphere_coefficient(n);
    c->light_coef = pow(s->light_unit_distance, n - 1);
    float lum = 0.2126f * s->background_rgb[0] + 0.7152f * s->background_rgb[1] + 0.0722f * s->background_rgb[2];
    c->bg_light[0] = c->bg_light[1] = c->bg_light[2] = lum;
    c->k = orc_samples_per_dimension(n - 1, s->samples_per_pixel);
    if (c->k < 1) return fail("bad sample count");
    c->nsamples = 1;
    for (int i = 0; i < n - 1; ++i) c->nsamples *= c->k;
    return 0;
}

/* The screen point of one (pixel, sample) in pixel units, before the
 * projector's -screen/2 shift: pixel coordinate (GlobalIndex order, x0
 * fastest, src/com/global_index.h) plus the sample's stratified jittered point
 * (StratifiedJitteredSampleEngine::generate, engine.h:86-156: sample index
 * digits in base k, dimension 0 least significant, offset digit/k plus a
 * uniform jitter in [0, 1/k)). */
static void sample_point(const orc_scene* s, int n, int k, int64_t pixel, int sample, double* out)
{
    int64_t rem = pixel;
    double u[8];
    uniforms4(s->seed, pixel, sample, s->pass_index, 0, 0, u);
    if (n - 1 > 4) uniforms4(s->seed, pixel, sample, s->pass_index, 0, 1, u + 4);
    int sr = sample;
    double rk = 1.0 / k;
    for (int i = 0; i < n - 1; ++i) {
        int64_t x = rem % s->screen[i];
        rem /= s->screen[i];
        int d = sr % k;
        sr /= k;
        double off = (double)d / k;
        double jitter = u[i] * rk;
        out[i] = (double)x + (off + jitter);
    }
}

int orc_sample_point(const orc_scene* s, int64_t pixel, int sample, double* point)
{
    int n = s->n;
    int k = orc_samples_per_dimension(n - 1, s->samples_per_pixel);
    if (k < 1) return fail("bad sample count");
    sample_point(s, n, k, pixel, sample, point);
    return 0;
}

/* paint_pixels body for one (pixel, sample), painter.cpp:419-442 */
static void sample_color(const pctx* c, pstate* st, float* rgb)
{
    const orc_scene* s = c->s;
    int n = c->m->n;
    double sp[NM - 1];
    sample_point(s, n, c->k, st->pixel, st->sample, sp);
    /* screen_org + point, projector.h:51-66 */
    for (int i = 0; i < n - 1; ++i) sp[i] = (-s->screen[i] * 0.5) + sp[i];
    double org[NM], dir[NM];
    projector_ray(s, n, c->cam_dir, (double (*)[NM])c->axes, c->square_radius, sp, org, dir);
    trace_path(c, st, 0, 1.0f, org, dir, 0, rgb);
}

void orc_trace_path_sample(const orc_mesh* m, const orc_scene* s, int64_t pixel, int sample, float* rgb,
                           int64_t* rays)
{
    pctx c;
    if (make_ctx(m, s, &c)) {
        rgb[0] = rgb[1] = rgb[2] = NAN;
        return;
    }
    pstate st = {pixel, sample, 0};
    sample_color(&c, &st, rgb);
    if (rays) *rays = st.rays;
}

typedef struct {
    const pctx* c;
    int64_t begin, count;
    const int64_t* list;  /* explicit pixel indices, or NULL: begin + j */
    int64_t* next;
    float* out;
    int64_t rays;
} worker;

static void* work(void* arg)
{
    worker* w = arg;
    const int64_t CH = 64;
    for (;;) {
        int64_t b = __atomic_fetch_add(w->next, CH, __ATOMIC_RELAXED);
        if (b >= w->count) break;
        int64_t e = b + CH < w->count ? b + CH : w->count;
        for (int64_t j = b; j < e; ++j) {
            pstate st = {w->list ? w->list[j] : w->begin + j, 0, 0};
            float sum[3] = {0, 0, 0};
            for (int s = 0; s < w->c->nsamples; ++s) {
                st.sample = s;
                float col[3];
                sample_color(w->c, &st, col);
                for (int k = 0; k < 3; ++k) sum[k] += col[k];
            }
            for (int k = 0; k < 3; ++k) w->out[j * 3 + k] = sum[k];
            w->rays += st.rays;
        }
    }
    return NULL;
}

static int render_impl(const orc_mesh* m, const orc_scene* s, int64_t pixel_begin, const int64_t* list,
                       int64_t pixel_count, int threads, float* rgb_sum, int64_t* rays)
{
    pctx c;
    if (make_ctx(m, s, &c)) return 1;
    if (threads < 1) return fail("Painter thread count (%d) must be greater than 0", threads);
    int64_t total_pixels = 1;
    for (int i = 0; i < s->n - 1; ++i) total_pixels *= s->screen[i];
    for (int64_t j = 0; list && j < pixel_count; ++j)
        if (list[j] < 0 || list[j] >= total_pixels) return fail("pixel %lld outside the screen", (long long)list[j]);
    int64_t next = 0;
    worker* ws = calloc(threads, sizeof(worker));
    pthread_t* th = calloc(threads, sizeof(pthread_t));
    for (int i = 0; i < threads; ++i) {
        ws[i].c = &c;
        ws[i].begin = pixel_begin;
        ws[i].list = list;
        ws[i].count = pixel_count;
        ws[i].next = &next;
        ws[i].out = rgb_sum;
        pthread_create(&th[i], NULL, work, &ws[i]);
    }
    int64_t total = 0;
    for (int i = 0; i < threads; ++i) {
        pthread_join(th[i], NULL);
        total += ws[i].rays;
    }
    free(ws);
    free(th);
    if (rays) *rays = total;
    return 0;
}

int orc_render(const orc_mesh* m, const orc_scene* s, int64_t pixel_begin, int64_t pixel_count, int threads,
               float* rgb_sum, int64_t* rays)
{
    return render_impl(m, s, pixel_begin, NULL, pixel_count, threads, rgb_sum, rays);
}

int orc_render_pixels(const orc_mesh* m, const orc_scene* s, const int64_t* pixels, int64_t count, int threads,
                      float* rgb_sum, int64_t* rays)
{
    return render_impl(m, s, 0, pixels, count, threads, rgb_sum, rays);
}
