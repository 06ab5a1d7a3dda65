// nsp_stages.hpp -- per-ray bodies of the wavefront stages (nsp_render.hip).
#pragma once

#include <cfloat>

#include "nsp_device.hpp"
#include "nsp_render.hpp"
#include "nsp_trace.hpp"

namespace nsp {

// paint_pixels' sample + Projector::ray (src/painter/painter.cpp:419-442,
// src/painter/sampling/engine.h:86-156, src/painter/projectors/projector.h:143-150
// (perspective), 194-201 (parallel), 259-274 (spherical)); Ray's constructor
// normalises the direction (src/com/ray.h:34-36)
template <int N, typename T>
__device__ __forceinline__ void primary_ray(const SceneK<T>& sc, int64_t pixel, int sample, T* org, T* dir)
{
    T u[8];
    uniforms4<T>(sc.seed, pixel, sample, pass_index(sc), 0, 0, u);
    if constexpr (N - 1 > 4) uniforms4<T>(sc.seed, pixel, sample, pass_index(sc), 0, 1, u + 4);
    int64_t rem = pixel;
    int sr = sample;
    T sp[N - 1];
#pragma unroll
    for (int i = 0; i < N - 1; ++i) {
        int64_t x = rem % sc.screen[i];
        rem /= sc.screen[i];
        int dgt = sr % sc.k;
        sr /= sc.k;
        T off = (T)dgt / (T)sc.k;
        T spt = off + u[i] * sc.k_inv;
        sp[i] = (-(T)sc.screen[i] * (T)0.5) + ((T)x + spt);
    }
    T sd[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        sd[j] = sc.axes[0][j] * sp[0];
#pragma unroll
        for (int i = 1; i < N - 1; ++i) sd[j] = sd[j] + sc.axes[i][j] * sp[i];
    }
    if (sc.proj == NSP_PROJECTOR_PARALLEL) {
#pragma unroll
        for (int j = 0; j < N; ++j) {
            org[j] = sc.cam_org[j] + sd[j];
            dir[j] = sc.cam_dir[j];
        }
        return;
    }
    T z = 1;
    if (sc.proj == NSP_PROJECTOR_SPHERICAL) {
        T r = sp[0] * sp[0];
#pragma unroll
        for (int i = 1; i < N - 1; ++i) r = fma(sp[i], sp[i], r);
        z = sqrt(sc.proj_r2 - r);  // positive: the scene constructor bounds the view angle
    }
#pragma unroll
    for (int j = 0; j < N; ++j) {
        org[j] = sc.cam_org[j];
        dir[j] = sc.cam_dir[j] * z + sd[j];
    }
    normalize<N, T>(dir);
}

// Parallelotope::intersect (parallelotope.h:196-246) and
// HyperplaneParallelotopeGeometry::intersect (hyperplane_geometry.h:41-52,
// 240-261), in FP64: the nearest t > 0 at which the ray meets the shape.
template <int N>
__device__ __forceinline__ bool shape_intersect(const ShapeRec& s, const double* o, const double* d, double* t)
{
    if (s.kind == NSP_SHAPE_PARALLELOTOPE) {
        double f_max = -DBL_MAX, b_min = DBL_MAX;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const double sd = dot<N, double>(d, s.pn[i]);
            const double od = dot<N, double>(o, s.pn[i]);
            if (sd == 0) {
                if (od - s.d1[i] > 0 || -od - s.d2[i] > 0) return false;  // parallel and outside
                continue;
            }
            const double a1 = (s.d1[i] - od) / sd;
            const double a2 = (s.d2[i] + od) / -sd;
            if (sd < 0) {
                f_max = fmax(a1, f_max);
                b_min = fmin(a2, b_min);
            } else {
                b_min = fmin(a1, b_min);
                f_max = fmax(a2, f_max);
            }
            if (b_min <= 0 || b_min < f_max) return false;
        }
        *t = f_max > 0 ? f_max : b_min;
        return true;
    }
    const double sd = dot<N, double>(s.n, d);
    if (sd == 0) return false;
    double v[N];
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = s.org[k] - o[k];
    const double tt = dot<N, double>(v, s.n) / sd;
    if (!(tt > 0)) return false;
    double p[N];
#pragma unroll
    for (int k = 0; k < N; ++k) p[k] = o[k] + d[k] * tt;
#pragma unroll
    for (int i = 0; i < N - 1; ++i) {
        const double c = dot<N, double>(p, s.pn[i]) - s.d1[i];
        if (c <= 0 || c >= 1) return false;
    }
    *t = tt;
    return true;
}

// Parallelotope::normal (parallelotope.h:261-290): the normal of the nearest
// face; HyperplaneParallelotope::normal: its plane's normal.
template <int N>
__device__ __forceinline__ void shape_normal(const ShapeRec& s, const double* p, double* n)
{
    if (s.kind != NSP_SHAPE_PARALLELOTOPE) {
#pragma unroll
        for (int k = 0; k < N; ++k) n[k] = s.n[k];
        return;
    }
    double mn = DBL_MAX;
    int best = 0;
    bool neg = false;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const double dd = dot<N, double>(p, s.pn[i]);
        double l = fabs(dd - s.d1[i]);
        if (l < mn) {
            mn = l;
            best = i;
            neg = false;
        }
        l = fabs(-dd - s.d2[i]);
        if (l < mn) {
            mn = l;
            best = i;
            neg = true;
        }
    }
#pragma unroll
    for (int k = 0; k < N; ++k) n[k] = neg ? -s.pn[best][k] : s.pn[best][k];
}

// The nearest shape hit before t_max (object order breaks exact ties), or -1.
template <int N, typename T>
__device__ __forceinline__ int nearest_shape(const SceneK<T>& sc, const T* o, const T* d, double t_max, double* t)
{
    double o64[N], d64[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        o64[k] = (double)o[k];
        d64[k] = (double)d[k];
    }
    int hit = -1;
    double best = t_max;
    for (int k = 0; k < sc.shape_count; ++k) {
        double ts;
        if (shape_intersect<N>(sc.shapes[k], o64, d64, &ts) && ts < best) {
            best = ts;
            hit = k;
        }
    }
    *t = best;
    return hit;
}

// light_source_is_visible over the shapes (painter.cpp:143-186): the first
// shape in object order that meets the ray before the light, or -1.
template <int N, typename T>
__device__ __forceinline__ int blocking_shape(const SceneK<T>& sc, const T* o, const T* d, double dist)
{
    double o64[N], d64[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        o64[k] = (double)o[k];
        d64[k] = (double)d[k];
    }
    for (int k = 0; k < sc.shape_count; ++k) {
        double ts;
        if (shape_intersect<N>(sc.shapes[k], o64, d64, &ts) && ts < dist) return k;
    }
    return -1;
}

template <int N, typename T>
__device__ __forceinline__ void load_normal(const MeshShadeArgs& M, int idx, T* gn)
{
    if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int k = 0; k < N; ++k) gn[k] = M.normal32[(int64_t)idx * N + k];
    } else {
#pragma unroll
        for (int k = 0; k < N; ++k) gn[k] = M.rec64[(int64_t)idx * rec64_stride(N) + k];
    }
}

// MeshHyperplaneSimplex::shading_normal, src/painter/shapes/mesh_hyperplane_simplex.cpp:157-186
template <int N, typename T>
__device__ void shading_normal(const MeshShadeArgs& M, int idx, const T* p, const T* gn, T* sn)
{
    int nt = M.ntype[idx];
    int type = nt & 3;
    if (type == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) sn[k] = gn[k];
        return;
    }
    const double* r = M.rec64 + (int64_t)idx * rec64_stride(N);
    double pd[N];
#pragma unroll
    for (int k = 0; k < N; ++k) pd[k] = p[k];
    double bc[N];
    bc[N - 1] = 1;
#pragma unroll
    for (int i = 0; i < N - 1; ++i) {
        bc[i] = dot<N, double>(pd, r + 2 * N + i * N) - r[2 * N + (N - 1) * N + i];
        bc[N - 1] -= bc[i];
    }
    double acc[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const double* vn = M.vn64 + (int64_t)M.fnrm[(int64_t)idx * N + k] * N;
        double sgn = (type == 2 && ((nt >> (8 + k)) & 1)) ? -1.0 : 1.0;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            double term = bc[k] * (sgn * vn[i]);
            acc[i] = (k == 0) ? term : acc[i] + term;
        }
    }
    normalize<N, double>(acc);
#pragma unroll
    for (int k = 0; k < N; ++k) sn[k] = (T)acc[k];
    normalize<N, T>(sn);
}

template <int N, typename T>
struct ShadeOut {
    float add[3];      // radiance added to the path now (background or emitter)
    bool next;         // a diffuse bounce ray continues the path
    T norg[N], ndir[N];
    float nbeta[3];
    float nlevel;
    bool shadow;       // a shadow ray towards the light
    T sorg[N], sdir[N];
    T stmax;
    float contrib[3];  // radiance added when the shadow ray is unoccluded
    int sshape;        // the first shape blocking the shadow ray, or -1
    float av[3 + N];   // first-hit albedo and normal (when asked for)
    // with several lights: the hit the shadow rays start from (direct: they are due)
    bool direct, mesh;
    T sp[N], sgn[N], ssn[N];
    float sbeta[3], sscol[3];
    // set by the caller: the shadow ray is kept as its hit point alone (sorg),
    // to be rebuilt by shadow_ray_from where it is tested
    bool spoint;
};

// The shadow ray from hit point p towards a point light at lpos: vl = lpos - p,
// ld = vl / |vl| its direction, dist = |vl| its length and o1 = p + ld * offset,
// from fused multiply-adds and, in FP32, one hardware reciprocal square root
// (v_rsq_f32, within 1 ulp; FP64: an IEEE square root and division), so every
// kernel that rebuilds the ray from p gets the same bits.
template <int N, typename T>
__device__ __forceinline__ void shadow_ray_from(const T* lpos, const T* p, T offset, T* vl, T* ld, T* dist, T* o1)
{
#pragma unroll
    for (int k = 0; k < N; ++k) vl[k] = lpos[k] - p[k];
    const T len2 = dot<N, T>(vl, vl);
    T inv;
    if constexpr (sizeof(T) == 4)
        inv = __builtin_amdgcn_rsqf(len2);
    else
        inv = (T)1 / sqrt(len2);
#pragma unroll
    for (int k = 0; k < N; ++k) ld[k] = vl[k] * inv;
    *dist = len2 * inv;
#pragma unroll
    for (int k = 0; k < N; ++k) o1[k] = fma(ld[k], offset, p[k]);
}

// The shadow ray of one light source from a hit (direct_diffuse_lighting's
// loop body, painter.cpp:205-289; PointLight and ConstantLight::properties,
// src/painter/lights/light_source.h:26-78): sets out.shadow and, when the light
// can reach the point, the ray, its distance limit, its contribution (path
// weight beta times surface colour scol times light colour and cosine weight)
// and the first shape in the way.
template <int N, typename T, bool SMOOTH, bool SHAPES>
__device__ __forceinline__ void light_shadow_ray(const SceneK<T>& sc, const TreeArgs& A, int l, const T* p,
                                                 const T* gn, const T* sn, bool mesh, const float* beta,
                                                 const float* scol, int2* st, unsigned long long* probe_count,
                                                 ShadeOut<N, T>& out)
{
    out.shadow = false;
    out.sshape = -1;
    T vl[N], ld[N], dist, o1[N];
    shadow_ray_from<N, T>(sc.light_pos[l], p, sc.ray_offset, vl, ld, &dist, o1);
    T sq2 = dot<N, T>(vl, vl);
    T scale;
    if (sc.light_kind[l] == 1)
        scale = (T)1;
    else if constexpr (N == 3)
        scale = sc.light_coef[l] / sq2;
    else
        scale = sc.light_coef[l] / pow(sq2, (T)(N - 1) / (T)2);
    float lc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) lc[c] = sc.light_rgb[l][c] * (float)scale;
    T dln = dot<N, T>(ld, sn);
    bool lit = !(fmaxf(lc[0], fmaxf(lc[1], lc[2])) < 1e-4f) && dln > (T)0;
    if (!lit) return;
    T tmaxs = dist;
    if (SMOOTH && mesh && sc.smooth && dot<N, T>(ld, gn) < (T)0) {
        // smooth-normal back face: forward and reverse probes (painter.cpp:238-284)
        atomicAdd(probe_count, 1ull);
        double o64[N], d64[N];
#pragma unroll
        for (int k = 0; k < N; ++k) {
            o64[k] = o1[k];
            d64[k] = ld[k];
        }
        HitResult h1 = traverse<N, T, false>(A, o1, ld, (T)INFINITY, st, o64, d64);
        T tf = (T)h1.t;
        lit = h1.idx >= 0 && tf < dist;
        if (lit) {
            atomicAdd(probe_count, 1ull);
            T rd[N], ro[N];
#pragma unroll
            for (int k = 0; k < N; ++k) {
                rd[k] = -ld[k];
                ro[k] = o1[k] + rd[k] * ((T)2 * sc.ray_offset);
                o64[k] = ro[k];
                d64[k] = rd[k];
            }
            HitResult h2 = traverse<N, T, false>(A, ro, rd, (T)INFINITY, st, o64, d64);
            lit = !(h2.idx >= 0 && (T)h2.t < tf);
#pragma unroll
            for (int k = 0; k < N; ++k) o1[k] = o1[k] + ld[k] * (tf + sc.ray_offset);
            tmaxs = dist - tf;
        }
    }
    if (!lit) return;
    float w = (float)(sc.cos_coef * dln);
    out.shadow = true;
    if (SHAPES && sc.shape_count > 0) out.sshape = blocking_shape<N, T>(sc, o1, ld, (double)dist);
#pragma unroll
    for (int k = 0; k < N; ++k) {
        out.sorg[k] = out.spoint ? p[k] : o1[k];
        out.sdir[k] = ld[k];
    }
    out.stmax = tmaxs;
#pragma unroll
    for (int c = 0; c < 3; ++c) out.contrib[c] = beta[c] * (scol[c] * (lc[c] * w));
}

// One path vertex: trace_path's shading, direct_diffuse_lighting's shadow-ray
// set-up and diffuse_lighting's bounce (src/painter/painter.cpp:197-397), for
// the ray (o, d) whose nearest hit is (idx, t).  With aov, out.av receives
// the first-hit albedo and normal.
// SHAPES: false compiles out the shapes beside the mesh (scenes without them).
// MULTI: the shadow rays are left to the caller, one light after another.
// SMOOTH: false compiles out the smooth-normal probes (scenes painted without
// smooth normals never take them).
template <int N, typename T, bool SMOOTH = true, bool SHAPES = true, bool MULTI = false>
__device__ __forceinline__ void shade_one(const SceneK<T>& sc, const TreeArgs& A, const MeshShadeArgs& M, int depth, int64_t pixel,
                          int sample, const T* o, const T* d, int idx, T t, const float* beta, float level, int2* st,
                          unsigned long long* probe_count, bool aov, ShadeOut<N, T>& out)
{
    out.add[0] = out.add[1] = out.add[2] = 0.0f;
    out.next = false;
    out.shadow = false;
    out.sshape = -1;
    out.direct = false;
    const bool diffuse_reflection = depth > 0;
    // ray_intersection over the objects (ray_intersection.h:20-80): the mesh
    // hit, or a nearer shape
    int shape = -1;
    if (SHAPES && sc.shape_count > 0) {
        double ts;
        shape = nearest_shape<N, T>(sc, o, d, idx >= 0 ? (double)t : DBL_MAX, &ts);
        if (shape >= 0) t = (T)ts;
    }
    const bool hit = idx >= 0 || shape >= 0;
    const bool mesh = shape < 0;
    T p[N];
#pragma unroll
    for (int k = 0; k < N; ++k) p[k] = o[k] + d[k] * t;
    // the surface properties of the hit (SurfaceProperties, src/painter/objects.h:27-139)
    T gn[N], sn[N];
    float rgb[3], light_rgb[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        rgb[c] = sc.obj_rgb[c];
        light_rgb[c] = sc.obj_light_rgb[c];
    }
    float diffuse = sc.diffuse;
    bool is_light = sc.obj_is_light;
    if (SHAPES && shape >= 0) {
        const ShapeRec& sr = sc.shapes[shape];
        double p64[N], n64[N];
#pragma unroll
        for (int k = 0; k < N; ++k) p64[k] = (double)p[k];
        shape_normal<N>(sr, p64, n64);
#pragma unroll
        for (int k = 0; k < N; ++k) gn[k] = (T)n64[k];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            rgb[c] = sr.rgb[c];
            light_rgb[c] = sr.light_rgb[c];
        }
        diffuse = sr.diffuse;
        is_light = sr.light;
    } else if (idx >= 0) {
        load_normal<N, T>(M, idx, gn);
    }
    if (aov) {
        T an[N];
        if (hit) {
#pragma unroll
            for (int k = 0; k < N; ++k) an[k] = gn[k];
            if (dot<N, T>(d, an) > (T)0)
#pragma unroll
                for (int k = 0; k < N; ++k) an[k] = -an[k];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) out.av[c] = hit ? diffuse * rgb[c] : sc.bg_rgb[c];
#pragma unroll
        for (int k = 0; k < N; ++k) out.av[3 + k] = hit ? (float)an[k] : 0.0f;
    }
    if (!hit) {
        // miss: background, a light source when reached by a diffuse bounce (painter.cpp:335-340)
        const float* bg = (sc.bg_is_light && diffuse_reflection) ? sc.bg_light_rgb : sc.bg_rgb;
#pragma unroll
        for (int c = 0; c < 3; ++c) out.add[c] = beta[c] * bg[c];
        return;
    }
    T dgn = dot<N, T>(d, gn);
    if (fabs(dgn) <= (T)0) return;
    if (is_light) {
#pragma unroll
        for (int c = 0; c < 3; ++c) out.add[c] = beta[c] * (diffuse_reflection ? light_rgb[c] : rgb[c]);
        return;
    }
    if (mesh && sc.smooth && sc.has_normals)
        shading_normal<N, T>(M, idx, p, gn, sn);
    else {
#pragma unroll
        for (int k = 0; k < N; ++k) sn[k] = gn[k];
    }
    if (dgn > 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            gn[k] = -gn[k];
            sn[k] = -sn[k];
        }
    }
    if (!(diffuse > 0)) return;
    float scol[3] = {diffuse * rgb[0], diffuse * rgb[1], diffuse * rgb[2]};
    float new_level = level * fmaxf(scol[0], fmaxf(scol[1], scol[2]));
    if (!(new_level >= 1e-4f)) return;
    // direct_diffuse_lighting, painter.cpp:197-292: one light here; several
    // lights are taken one after another by the caller from the hit in out
    if (MULTI) {
        out.direct = true;
        out.mesh = mesh;
#pragma unroll
        for (int k = 0; k < N; ++k) {
            out.sp[k] = p[k];
            out.sgn[k] = gn[k];
            out.ssn[k] = sn[k];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            out.sbeta[c] = beta[c];
            out.sscol[c] = scol[c];
        }
    } else if (sc.light_count > 0) {
        light_shadow_ray<N, T, SMOOTH, SHAPES>(sc, A, 0, p, gn, sn, mesh, beta, scol, st, probe_count, out);
    }
    // diffuse_lighting, painter.cpp:298-323
    if (depth < sc.max_depth) {
        T u[8];
        uniforms4<T>(sc.seed, pixel, sample, pass_index(sc), depth, 2, u);
        if constexpr (N >= 6) uniforms4<T>(sc.seed, pixel, sample, pass_index(sc), depth, 3, u + 4);
        T nd[N];
        cosine_hemisphere<N, T>(sn, u, nd);
        normalize<N, T>(nd);
        // only a mesh's bounce must leave its geometric side (painter.cpp:308-313)
        if (!mesh || dot<N, T>(nd, gn) > (T)0) {
            out.next = true;
#pragma unroll
            for (int k = 0; k < N; ++k) {
                out.norg[k] = p[k] + nd[k] * sc.ray_offset;
                out.ndir[k] = nd[k];
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) out.nbeta[c] = beta[c] * scol[c];
            out.nlevel = new_level;
        }
    }
}

}  // namespace nsp
