// nsp_trace.hpp -- 2^N-tree traversal for FP32 (with FP64 near-boundary
// re-test) and FP64 rays, shared by the first-hit kernel and the wavefront
// render stages.
//
// Replaces SpatialSubdivisionTree::trace_ray (src/painter/space/tree.h:534-590),
// which re-descends from the root for every box it steps through and accepts a
// leaf's nearest simplex only inside that leaf box.  Here the tree is a
// hierarchy of regular grids, so a point's cell at every level follows from
// its integer coordinates at the finest level Q = floor((p - lo) / cell_min).
// Each lane keeps the node records (int2: first child and stride, or entry
// range) of its current root-to-leaf path in an LDS stack, one per level;
// after leaving a leaf it pops only the levels whose cell differs (bit length
// of the XOR of old and new coordinates) and descends from the deepest kept
// record, so a restart reads its first node from LDS.  Leaves are visited in
// ray order and the walk stops once the next leaf starts beyond the nearest
// hit, so the result is the nearest hit over all simplices with the
// reference's tie rule (smaller index first,
// src/painter/space/ray_intersection.h:141-150).
#pragma once

#include <type_traits>

#include "nsp_device.hpp"
#include "nsp_internal.hpp"

namespace nsp {

#ifndef NSP_TRACE_BLOCK
#define NSP_TRACE_BLOCK 256
#endif
constexpr int TRACE_BLOCK = NSP_TRACE_BLOCK;  // threads per block of the trace and stage kernels
// Levels (the root's included) of a walk's LDS node stack: the tree depth
// limit per dimension (src/painter/shapes/mesh.cpp:33-58; the tree build never
// makes more levels), so that an n = 4 or 5 walk keeps 8 or 6 records per lane
// instead of 10.
__host__ __device__ constexpr int stack_levels(int n) { return n <= 3 ? 10 : n == 4 ? 8 : n == 5 ? 6 : 5; }

// Half-width of the barycentric band (in barycentric units, scaled by
// 1 + 1/|cos|) inside which an FP32 simplex decision is re-taken in FP64, and
// the |cos| below which an FP32 hit is re-taken in FP64 (near-parallel).
constexpr float RETEST_BAND = 2.0e-4f;
constexpr float PARALLEL_COS = 0.05f;

struct TreeArgs {
    const int2* nodes;
    const int32_t* entries;
    const float* soa;
    int64_t pad;
    const float* rec32;  // simplex-ordered AoS records (layout 1)
    int layout;          // 0: leaf-ordered SoA, 1: simplex-ordered AoS gathered through the entry ids
    const double* rec64;
    float lo[NMAX], size[NMAX], inv_fine[NMAX];
    double lo64[NMAX], size64[NMAX], inv_fine64[NMAX];
    int lmax;
    float tadv;
    double tadv64;
    unsigned long long* counters;  // node visits, leaf visits, simplex tests, retests (or null)
};

template <typename T>
struct TreeBox;
template <>
struct TreeBox<float> {
    __device__ static const float* lo(const TreeArgs& a) { return a.lo; }
    __device__ static const float* size(const TreeArgs& a) { return a.size; }
    __device__ static const float* inv(const TreeArgs& a) { return a.inv_fine; }
    __device__ static float adv(const TreeArgs& a) { return a.tadv; }
};
template <>
struct TreeBox<double> {
    __device__ static const double* lo(const TreeArgs& a) { return a.lo64; }
    __device__ static const double* size(const TreeArgs& a) { return a.size64; }
    __device__ static const double* inv(const TreeArgs& a) { return a.inv_fine64; }
    __device__ static double adv(const TreeArgs& a) { return a.tadv64; }
};

// FP64 simplex test exactly as the reference states it:
// HyperplaneSimplexGeometry::intersect + plane_intersect
// (src/painter/space/hyperplane_geometry.h:40-52, :181-204), no fused ops.
template <int N>
__device__ __forceinline__ bool simplex_test64(const double* __restrict__ r, const double* o,
                                                         const double* d, double* t_out)
{
#pragma clang fp contract(off)
    const double* nrm = r;
    const double* v0 = r + N;
    double s = dot<N, double>(nrm, d);
    if (s == 0) return false;
    double diff[N];
#pragma unroll
    for (int i = 0; i < N; ++i) diff[i] = v0[i] - o[i];
    double t = dot<N, double>(diff, nrm) / s;
    if (!(t > 0)) return false;
    double p[N];
#pragma unroll
    for (int i = 0; i < N; ++i) p[i] = o[i] + d[i] * t;
    double last = 1;
    bool ok = true;
#pragma unroll
    for (int i = 0; i < N - 1; ++i) {
        double b = dot<N, double>(p, r + 2 * N + i * N) - r[2 * N + (N - 1) * N + i];
        if (b <= 0 || b >= 1) ok = false;
        last -= b;
    }
    if (!(last > 0)) ok = false;
    *t_out = t;
    return ok;
}

struct HitResult {
    int idx;
    double t;
    bool near;
};

template <int N>
__device__ __forceinline__ void load_vec_f(const float* __restrict__ soa, int64_t pad, int64_t e, int c0, float* v)
{
#pragma unroll
    for (int q = 0; q < N; ++q) v[q] = soa[(int64_t)(c0 + q) * pad + e];
}

// Results of a candidate test.
constexpr int CAND_MISS = 0, CAND_HIT = 1, CAND_DEFER = 2;

// The order of hits along a ray: the smaller distance, then the smaller
// simplex index.  FP32 walks compare distances rounded to FP32 (the distance a
// hit keeps), so that every walk and the rasteriser, which orders hits by
// (FP32 distance, index) keys, choose the same simplex among hits whose
// distances round alike, whatever order they meet them in; FP64 walks compare
// FP64 distances.  A candidate test accepts a hit at the limit distance when
// its index is below best_idx, the index of the hit holding the limit.
template <typename T>
__device__ __forceinline__ bool nearer(double t, int sid, double best_t, int best)
{
    if constexpr (sizeof(T) == 4) {
        const float a = (float)t, b = (float)best_t;
        return a < b || (a == b && sid < best);
    } else {
        return t < best_t || (t == best_t && sid < best);
    }
}

// The FP32 test of one simplex against the ray (o, d): comp(c) returns
// component c of the simplex's FP32 record (normal[N], d, then the N-1
// barycentric planes), sid() its index.  Shared by the tree walk and the
// primary-visibility rasteriser, so both make the same decisions.
template <int N, bool DEFER, class Comp, class Sid>
__device__ __forceinline__ int fp32_test(const TreeArgs& A, const Comp& comp, const Sid& get_sid, const float* o,
                                         const float* d, const double* o64, const double* d64, float tlimit,
                                         int best_idx, double* t_hit, int* idx, bool* near, unsigned* retests)
{
    float nrm[N];
#pragma unroll
    for (int q = 0; q < N; ++q) nrm[q] = comp(q);
    float s = dot<N, float>(nrm, d);
    if (s == 0.0f) return CAND_MISS;
    float dd = comp(N);
    float no = dot<N, float>(nrm, o);
    // one hardware reciprocal (v_rcp_f32, within 1 ulp) instead of three
    // IEEE divisions of about ten instructions each
    const float r = __builtin_amdgcn_rcpf(s);
    const float ar = fabsf(r);
    float t = (dd - no) * r;
    float as = fabsf(s);
    float band = RETEST_BAND * (1.0f + ar);
    // rounding bound of t (the numerator's cancellation, the reciprocal's
    // ulp and the product's rounding); outside it the sign of t and the
    // comparison with the limit are decided in FP32
    float terr = 6.0e-7f * (fabsf(dd) + fabsf(no)) * ar;
    if (t < -terr) return CAND_MISS;
    if (t > tlimit + terr) return CAND_MISS;
    float p[N];
#pragma unroll
    for (int i = 0; i < N; ++i) p[i] = o[i] + d[i] * t;
    float last = 1.0f, margin = 3.0e38f;
    bool in = true;
#pragma unroll
    for (int i = 0; i < N - 1; ++i) {
        float pn[N];
#pragma unroll
        for (int q = 0; q < N; ++q) pn[q] = comp(N + 1 + i * (N + 1) + q);
        float b = dot<N, float>(p, pn) - comp(N + 1 + i * (N + 1) + N);
        in = in && (b > 0.0f) && (b < 1.0f);
        last -= b;
        margin = fminf(margin, b);
        if (b < -band) return CAND_MISS;  // decided in FP32: outside by more than the band
    }
    in = in && (last > 0.0f);
    margin = fminf(margin, last);
    bool retest = fabsf(margin) < band || (in && as < PARALLEL_COS) || t < terr;
    if (!retest) {
        if (!in || !(t > 0.0f)) return CAND_MISS;
        const int sid = get_sid();
        if (fabsf(t - tlimit) < terr && sid != best_idx) *near = true;  // near-tie with another simplex
        if (!(t < tlimit || (t == tlimit && sid < best_idx))) return CAND_MISS;
        *t_hit = (double)t;
        *idx = sid;
        return CAND_HIT;
    }
    if constexpr (DEFER) return CAND_DEFER;
    ++*retests;
    *near = true;
    const int sidx = get_sid();
    // without a separate FP64 ray the re-test uses the FP32 ray exactly
    double ob[N], db[N];
    if (!o64) {
#pragma unroll
        for (int q = 0; q < N; ++q) {
            ob[q] = o[q];
            db[q] = d[q];
        }
        o64 = ob;
        d64 = db;
    }
    double t64;
    if (!simplex_test64<N>(A.rec64 + (int64_t)sidx * rec64_stride(N), o64, d64, &t64)) return CAND_MISS;
    const float tf = (float)t64;
    if (!(tf < tlimit || (tf == tlimit && sidx < best_idx))) return CAND_MISS;
    *t_hit = t64;
    *idx = sidx;
    return CAND_HIT;
}

// Candidate test of one leaf entry: CAND_HIT for a hit with 0 < t < tlimit,
// CAND_MISS otherwise.  best_idx: the simplex that currently holds tlimit (a
// simplex listed in several leaves meets itself again and is no tie).
// LAY: the FP32 record layout (0 leaf-ordered SoA, 1 simplex-ordered AoS,
// -1 read from A.layout).  DEFER: instead of the FP64 re-test, return
// CAND_DEFER (the caller decides the candidate later).
// The simplex-ordered FP32 record of simplex sid, read with 16-byte loads.
template <int N>
__device__ __forceinline__ void load_rec32(const TreeArgs& A, int sid, float* rec)
{
    constexpr int RS = rec32_stride(N);
    const float4* r4 = reinterpret_cast<const float4*>(A.rec32 + (int64_t)sid * RS);
#pragma unroll
    for (int k = 0; k < RS / 4; ++k) {
        const float4 v = r4[k];
        rec[4 * k] = v.x;
        rec[4 * k + 1] = v.y;
        rec[4 * k + 2] = v.z;
        rec[4 * k + 3] = v.w;
    }
}

// The same record with every component in registers at return: the test's
// early exits would otherwise let the compiler sink the later loads into
// them, one memory round trip per exit (the normal, then d, then the planes).
template <int N>
__device__ __forceinline__ void load_rec32_now(const TreeArgs& A, int sid, float* rec)
{
    constexpr int RS = rec32_stride(N);
    const float4* r4 = reinterpret_cast<const float4*>(A.rec32 + (int64_t)sid * RS);
    float4 v[RS / 4];
#pragma unroll
    for (int k = 0; k < RS / 4; ++k) v[k] = r4[k];
#pragma unroll
    for (int k = 0; k < RS / 4; ++k) {
        asm volatile("" ::"v"(v[k].x), "v"(v[k].y), "v"(v[k].z), "v"(v[k].w));
        rec[4 * k] = v[k].x;
        rec[4 * k + 1] = v[k].y;
        rec[4 * k + 2] = v[k].z;
        rec[4 * k + 3] = v[k].w;
    }
}

template <int N, typename T, int LAY = -1, bool DEFER = false>
__device__ __forceinline__ int candidate(const TreeArgs& A, int64_t e, const T* o, const T* d, const double* o64,
                                         const double* d64, T tlimit, int best_idx, double* t_hit, int* idx,
                                         bool* near, unsigned* retests)
{
    if constexpr (sizeof(T) == 8) {
        int s = A.entries[e];
        double t;
        if (!simplex_test64<N>(A.rec64 + (int64_t)s * rec64_stride(N), o, d, &t)) return CAND_MISS;
        if (!(t < tlimit || (t == tlimit && s < best_idx))) return CAND_MISS;
        *t_hit = t;
        *idx = s;
        return CAND_HIT;
    } else {
        // component c of the candidate's FP32 record: leaf-ordered SoA (layout 0)
        // or the simplex-ordered AoS record read with 16-byte loads (layout 1)
        constexpr int RS = rec32_stride(N);
        float rec[RS];
        int sid = -1;
        const bool aos = LAY < 0 ? A.layout == 1 : LAY == 1;
        if (aos) {
            sid = A.entries[e];
            load_rec32<N>(A, sid, rec);
        }
        const float* soa = A.soa;
        const int64_t pad = A.pad;
        auto comp = [&](int c) -> float { return aos ? rec[c] : soa[(int64_t)c * pad + e]; };
        auto get_sid = [&]() -> int { return sid >= 0 ? sid : A.entries[e]; };
        return fp32_test<N, DEFER>(A, comp, get_sid, o, d, o64, d64, tlimit, best_idx, t_hit, idx, near, retests);
    }
}

// Slab interval of the axis-aligned box [blo, blo + bsz] (ParallelotopeOrtho
// semantics, src/painter/space/parallelotope_ortho.h:254-313, incl. s == 0).
template <int N, typename T>
__device__ __forceinline__ bool slab(const T* blo, const T* bsz, const T* o, const T* inv, const T* d, T* tn, T* tf)
{
    T a = -INFINITY, b = INFINITY;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        T lo = blo[i], hi = blo[i] + bsz[i];
        if (d[i] == 0) {
            if (o[i] < lo || o[i] > hi) return false;
            continue;
        }
        T x0 = (lo - o[i]) * inv[i], x1 = (hi - o[i]) * inv[i];
        a = fmax(a, fmin(x0, x1));
        b = fmin(b, fmax(x0, x1));
    }
    *tn = a;
    *tf = b;
    return a <= b;
}

// Exit of the ray from the located leaf box [blo, blo + bsz].  Axes along
// which the ray does not move are left out: the leaf was located from the
// ray's own coordinate there, so the ray stays in it along those axes even
// when that coordinate lies within rounding of a cell boundary and outside the
// box's computed bounds.  (Testing them made such an axis-parallel ray creep
// through its leaves one advance step at a time until the step limit ended
// the walk without its hit.)  An exit at or before t_cur (rounding along a
// moving axis) makes advance() double its step.
template <int N, typename T>
__device__ __forceinline__ T box_exit(const T* blo, const T* bsz, const T* o, const T* inv, const T* d)
{
    T b = INFINITY;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        if (d[i] == 0) continue;
        const T x0 = (blo[i] - o[i]) * inv[i], x1 = (blo[i] + bsz[i] - o[i]) * inv[i];
        b = fmin(b, fmax(x0, x1));
    }
    return b;
}

// The tree walk of one ray, split into start() and step() so that a lane can
// hold it across the iterations of a persistent kernel.  Nearest hit
// (ANY=false) or any hit before tmax (ANY=true).  The walk locates the leaf
// that contains the point t_cur + adv (integer cell coordinates at the finest
// level; the per-level node stack in LDS is kept down to the deepest common
// ancestor of the previous leaf), tests the leaf's candidates, and moves t_cur
// to the leaf box's exit.  It visits leaves in ray order and keeps the global
// nearest hit; it ends at the root box's exit or once the nearest hit lies
// before the current leaf.
//
// LAY: as for candidate().  DEFER: a candidate within the FP64 re-test band
// is kept as pending (up to PENDING per walk) and the walk goes on with the
// FP32 decisions only; the caller decides the pending candidates in FP64
// afterwards (pending_resolve) against the walk's nearest hit.  The nearest
// hit is the same as with the re-test inline: a pending candidate would only
// have lowered the limit later candidates are tested against, and the walk
// ends no earlier for it.  A walk with more pending candidates ends with
// deferred set, for a re-walk with the re-test inline.
// (C1, one MI355X: 4 pending candidates ran at 2347 Mrays/s, 8 at 2445, 12 at
// 2459; a walk that overflows its list is walked again with the FP64 re-test
// inline, a slow single-lane tail after each walk launch.)
#ifndef NSP_PENDING
#define NSP_PENDING 8
#endif
constexpr int PENDING = NSP_PENDING;

template <int N, typename T, bool ANY, int LAY = -1, bool DEFER = false>
struct RayWalk {
    // FP32 walks with deferred FP64 decisions find FP32 distances only
    using BestT = typename std::conditional<DEFER && sizeof(T) == 4, float, double>::type;
    T o[N], d[N], inv[N];
    T t_cur, t_end, adv, tmax;
    T t_next;  // exit of the located leaf box
    BestT best_t;
    int best, level, last_leaf;
    // finest-level cell coordinates of the last located point, lmax bits each
    // (n * lmax <= 28 for every supported dimension)
    unsigned qp;
    int pend[PENDING];
    int npend;
    unsigned steps;
    bool near, deferred;
    unsigned nvis, lvis, ntest, nret;

    // Enters the root box with o and d set; false when the ray misses it.
    __device__ __forceinline__ bool start(const TreeArgs& A, T tmax_)
    {
        tmax = tmax_;
        best_t = (BestT)tmax;
        best = -1;
        level = -1;
        last_leaf = -1;
        steps = 0;
        near = false;
        deferred = false;
        npend = 0;
        nvis = lvis = ntest = nret = 0;
#pragma unroll
        for (int i = 0; i < N; ++i) inv[i] = (T)1 / d[i];
        T t_in, t_out;
        if (!slab<N, T>(TreeBox<T>::lo(A), TreeBox<T>::size(A), o, inv, d, &t_in, &t_out)) return false;
        if (t_out < 0) return false;
        t_cur = fmax(t_in, (T)0);
        t_end = fmin(t_out, tmax);
        adv = TreeBox<T>::adv(A);
        return true;
    }

    // The first part of a step: locates the leaf that holds the next point of
    // the ray and its box exit (t_next).  True when the walk has ended;
    // otherwise *cnt candidates from entry *e0 are to be tested (0: an empty
    // leaf, or the leaf just tested).
    // stack: this lane's LDS stack, stack[l * TRACE_BLOCK] = the node record
    // of the walk's node at level l, so that a restart after a pop reads the
    // node from LDS instead of global memory.
    __device__ __forceinline__ bool locate(const TreeArgs& A, int2* stack, int* e0, int* cnt)
    {
        const T* rlo = TreeBox<T>::lo(A);
        const T* rsz = TreeBox<T>::size(A);
        const T* rinv = TreeBox<T>::inv(A);
        const T adv0 = TreeBox<T>::adv(A);
        const int lmax = A.lmax;
        const int R = 1 << lmax;
        if (++steps > 65536) return true;
        // the walk ends at the root box's exit: with t_cur == t_end the next
        // point lies outside the root and clamps back into the last cell
        if (t_cur >= t_end) return true;
        if (best >= 0 && (T)best_t < t_cur - adv0 * 4) return true;
        T tp = t_cur + adv;
        int Q[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            T x = floor((o[i] + d[i] * tp - rlo[i]) * rinv[i]);
            Q[i] = (int)fmin(fmax(x, (T)0), (T)(R - 1));
        }
        int2 nd;
        if (level < 0) {
            level = 0;
            nd = A.nodes[0];
            ++nvis;
            stack[0] = nd;
        } else {
            unsigned diff = 0;
            const unsigned cmask = (1u << lmax) - 1u;
#pragma unroll
            for (int i = 0; i < N; ++i)
                diff |= ((unsigned)Q[i] ^ ((qp >> (i * lmax)) & cmask)) >> (lmax - level);
            int pops = diff ? 32 - __clz(diff) : 0;
            level -= pops;
            nd = stack[level * TRACE_BLOCK];
        }
        while (nd.y < 0) {
            int shift = lmax - level - 1;
            int code = 0;
#pragma unroll
            for (int i = 0; i < N; ++i) code |= ((Q[i] >> shift) & 1) << i;
            const int node = nd.x + code * (-nd.y);
            ++level;
            nd = A.nodes[node];
            ++nvis;
            stack[level * TRACE_BLOCK] = nd;
        }
        qp = 0;
#pragma unroll
        for (int i = 0; i < N; ++i) qp |= (unsigned)Q[i] << (i * lmax);
        // leaf box interval
        T blo[N], bsz[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            T cell = ldexp(rsz[i], -level);
            bsz[i] = cell;
            blo[i] = rlo[i] + (T)(Q[i] >> (lmax - level)) * cell;
        }
        t_next = box_exit<N, T>(blo, bsz, o, inv, d);
        *cnt = 0;
        if (nd.x != last_leaf && nd.y > 0) {
            ++lvis;
            *e0 = nd.x;
            *cnt = nd.y;
        }
        if (nd.y > 0) last_leaf = nd.x;
        return false;
    }

    // The last part of a step: moves t_cur to the located leaf's exit.  True
    // when the walk has ended.
    __device__ __forceinline__ bool advance(const TreeArgs& A)
    {
        const T adv0 = TreeBox<T>::adv(A);
        if (t_next > t_cur) {
            t_cur = t_next;
            adv = adv0;
        } else {
            adv = adv * 2;  // no progress: the located cell ended before t_cur (rounding)
            if (t_cur + adv > t_end) return true;
        }
        return false;
    }

    // One leaf of the walk; true when the walk has ended.
    // o64/d64: the ray in FP64 for the near-boundary re-test, or null to use
    // the FP32 ray converted exactly.
    __device__ __forceinline__ bool step(const TreeArgs& A, int2* stack, const double* o64, const double* d64)
    {
        int e_begin = 0, e_count = 0;
        if (locate(A, stack, &e_begin, &e_count)) return true;
        if (e_count > 0) {
            // outcome of one candidate test; true ends the walk
            auto take = [&](int c, double th, int si, bool nr, int sid) -> bool {
                near = near || nr;
                if (DEFER && c == CAND_DEFER) {
                    near = true;
                    ++nret;  // decided in FP64 after the walk
                    bool known = false;
#pragma unroll
                    for (int k = 0; k < PENDING; ++k) known = known || (k < npend && pend[k] == sid);
                    if (!known) {
                        if (npend == PENDING) {
                            deferred = true;
                            return true;
                        }
#pragma unroll
                        for (int k = 0; k < PENDING; ++k)
                            if (k == npend) pend[k] = sid;
                        ++npend;
                    }
                    return false;
                }
                if (c == CAND_HIT) {
                    if (ANY) {
                        best_t = (BestT)th;
                        best = si;
                        return true;
                    }
                    if (nearer<T>(th, si, best_t, best)) {
                        best_t = (BestT)th;
                        best = si;
                    }
                }
                return false;
            };
            const int64_t e1 = (int64_t)e_begin + e_count;
            if constexpr (LAY == 1 && sizeof(T) == 4) {
                // two candidates at a time: both entry ids, then both records, are
                // loaded before either is tested (two independent load chains)
                constexpr int RS = rec32_stride(N);
                for (int64_t e = e_begin; e < e1; e += 2) {
                    const bool two = e + 1 < e1;
                    const int sa = A.entries[e];
                    const int sb = two ? A.entries[e + 1] : sa;
                    float ra[RS], rb[RS];
                    load_rec32<N>(A, sa, ra);
                    if (two) load_rec32<N>(A, sb, rb);
#pragma unroll
                    for (int g = 0; g < 2; ++g) {
                        if (g == 1 && !two) break;
                        const int sid = g ? sb : sa;
                        auto comp = [&](int cc) -> float { return g ? rb[cc] : ra[cc]; };
                        auto get_sid = [&]() -> int { return sid; };
                        ++ntest;
                        double th;
                        int si;
                        bool nr = false;
                        const int c = fp32_test<N, DEFER>(A, comp, get_sid, o, d, o64, d64, ANY ? tmax : (T)best_t,
                                                          best, &th, &si, &nr, &nret);
                        if (take(c, th, si, nr, sid)) return true;
                    }
                }
            } else {
                for (int64_t e = e_begin; e < e1; ++e) {
                    ++ntest;
                    double th;
                    int si;
                    bool nr = false;
                    T lim = ANY ? tmax : (T)best_t;
                    const int c = candidate<N, T, LAY, DEFER>(A, e, o, d, o64, d64, lim, best, &th, &si, &nr, &nret);
                    if (take(c, th, si, nr, DEFER && c == CAND_DEFER ? A.entries[e] : -1)) return true;
                }
            }
        }
        return advance(A);
    }

    __device__ __forceinline__ HitResult result() const
    {
        HitResult r{-1, 0.0, near};
        if (best >= 0) {
            r.idx = best;
            r.t = best_t;
        }
        return r;
    }

    __device__ __forceinline__ void flush_counters(const TreeArgs& A) const { flush_to(A.counters); }

    // adds the walk's traversal counters to a lane's running sums acc[0..3]
    __device__ __forceinline__ void add_to(unsigned long long* acc) const
    {
        acc[0] += nvis;
        acc[1] += lvis;
        acc[2] += ntest;
        acc[3] += nret;
    }

    // adds the walk's traversal counters to c[0..3] (node visits, leaf visits,
    // simplex tests, FP64 decisions) when c is not null
    __device__ __forceinline__ void flush_to(unsigned long long* c) const
    {
        if (c) {
            atomicAdd(&c[0], (unsigned long long)nvis);
            atomicAdd(&c[1], (unsigned long long)lvis);
            atomicAdd(&c[2], (unsigned long long)ntest);
            atomicAdd(&c[3], (unsigned long long)nret);
        }
    }
};

// LDS of one wave for coop_test, and the pending lists of its lanes' walks.
struct alignas(16) CoopWave {
    int incl[64];                 // inclusive count of candidates up to each lane
    int incl8[8];                 // incl of lanes 7, 15, ..., 63 (the owner search's first level)
    int ebase[64];                // per lane: the entry of its candidate list's slot 0 (e0 - exclusive count)
    unsigned long long key[64];   // per lane: nearest hit (t bits << 32 | simplex), or any hit
    int npend[64];                // per lane: its walk's pending candidates (more than PENDING: overflowed)
    int ndefer[64];               // per lane: deferred decisions of the current leaf
    int pend[64 * PENDING];       // per lane: its walk's pending candidates (the walk's list lives here)
};

// Sort key of a hit: t > 0, so its float bits order as unsigned integers; ties
// go to the smaller simplex index (the walk's tie rule).
__device__ __forceinline__ unsigned long long hit_key(float t, int sid)
{
    return ((unsigned long long)__float_as_uint(t) << 32) | (unsigned)sid;
}
constexpr unsigned long long NO_HIT_KEY = (0x7f800000ull << 32) | 0xffffffffull;

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Wave-cooperative FP32 test of the leaf candidates of every lane whose walk
// located a leaf (active, cnt candidates from entry e0; the FP32 AoS records,
// deferred FP64 decisions).  Called by all 64 lanes together; each lane's
// any_hit says which kind its walk is (a wave may hold both kinds).  The lanes'
// candidate lists are concatenated (a prefix sum over the wave) and dealt out
// 64 at a time: lane j of a round tests candidate base + j for its owner, the
// lane whose list holds it, with the owner's ray and limit read across lanes.
// A walk with one candidate and a walk with forty then cost the wave the same
// rounds, where testing each lane's own list keeps the wave until its longest
// list is done.
//   nearest hit: the owner's key is lowered with an LDS atomic min from its
//     hit so far, which keeps the nearest hit and, at equal t, the smaller
//     index.  Every candidate is tested against the limit the leaf started
//     with; a hit can only lower it, so the minimum is the one the sequential
//     walk finds.
//   any hit: any hit before tmax ends the owner's walk.
//   deferred candidates join the owner's pending list in L (no duplicates);
//     more than PENDING set deferred, for a re-walk.  The list stays in LDS for
//     the whole walk: the caller empties it (L.npend[lane] = 0) when the lane
//     starts a walk and reads it when the walk ends.
template <int N, class W>
__device__ __forceinline__ void coop_test(const TreeArgs& A, CoopWave& L, W& w, bool active, int e0, int cnt,
                                          bool any)
{
    constexpr int RS = rec32_stride(N);
    const int lane = threadIdx.x & 63;
    const int c = active ? cnt : 0;
    int incl = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    const int total = __shfl(incl, 63);
    if (total == 0) return;
    const float tl = any ? (float)w.tmax : (float)w.best_t;
    const int np0 = L.npend[lane];  // at most PENDING: a walk whose list overflowed has ended
    L.incl[lane] = incl;
    if ((lane & 7) == 7) L.incl8[lane >> 3] = incl;
    L.ebase[lane] = e0 - (incl - c);
    L.key[lane] = any || w.best < 0 ? NO_HIT_KEY : hit_key((float)w.best_t, w.best);
    L.ndefer[lane] = 0;
    wave_lds_sync();
    // owner of slot jj: the first lane whose inclusive count exceeds jj, i.e.
    // the number of lanes whose count is at most jj (the counts never fall):
    // whole groups of 8 lanes first, then the lanes of the next group
    auto owner_of = [&](int jj) -> int {
        const int4 g0 = *reinterpret_cast<const int4*>(&L.incl8[0]);
        const int4 g1 = *reinterpret_cast<const int4*>(&L.incl8[4]);
        const int g = (g0.x <= jj) + (g0.y <= jj) + (g0.z <= jj) + (g0.w <= jj) + (g1.x <= jj) + (g1.y <= jj) +
                      (g1.z <= jj) + (g1.w <= jj);
        const int4 l0 = *reinterpret_cast<const int4*>(&L.incl[8 * g]);
        const int4 l1 = *reinterpret_cast<const int4*>(&L.incl[8 * g + 4]);
        return 8 * g + (l0.x <= jj) + (l0.y <= jj) + (l0.z <= jj) + (l0.w <= jj) + (l1.x <= jj) + (l1.y <= jj) +
               (l1.z <= jj) + (l1.w <= jj);
    };
    // each round's owners and entry ids are found during the round before, so
    // the entry load of the next round travels with this round's record load
    int owner_n = owner_of(lane < total ? lane : total - 1);
    int sid_n = A.entries[L.ebase[owner_n] + (lane < total ? lane : total - 1)];
    for (int base = 0; base < total; base += 64) {
        const int j = base + lane;
        const int owner = owner_n;
        const int sid = sid_n;
        if (base + 64 < total) {
            const int jn = base + 64 + lane < total ? base + 64 + lane : total - 1;
            owner_n = owner_of(jn);
            sid_n = A.entries[L.ebase[owner_n] + jn];
        }
        float oo[N], od[N];
#pragma unroll
        for (int k = 0; k < N; ++k) {
            oo[k] = __shfl((float)w.o[k], owner);
            od[k] = __shfl((float)w.d[k], owner);
        }
        const float tlo = __shfl(tl, owner);
        const int bo = __shfl(any ? -1 : w.best, owner);
        const int npo = __shfl(np0, owner);
        if (j < total) {
            float rec[RS];
            load_rec32_now<N>(A, sid, rec);
            auto comp = [&](int cc) -> float { return rec[cc]; };
            auto get_sid = [&]() -> int { return sid; };
            double th;
            int si;
            bool nr = false;
            unsigned dummy = 0;
            const int r = fp32_test<N, true>(A, comp, get_sid, oo, od, nullptr, nullptr, tlo, bo, &th, &si, &nr,
                                             &dummy);
            if (r == CAND_HIT) {
                // (an any-hit owner only asks whether its key fell below NO_HIT_KEY)
                atomicMin(&L.key[owner], hit_key((float)th, si));
            } else if (r == CAND_DEFER) {
                atomicAdd(&L.ndefer[owner], 1);
                bool known = false;
#pragma unroll
                for (int k = 0; k < PENDING; ++k) known = known || (k < npo && L.pend[owner * PENDING + k] == sid);
                if (!known) {
                    const int slot = atomicAdd(&L.npend[owner], 1);
                    if (slot < PENDING) L.pend[owner * PENDING + slot] = sid;
                }
            }
        }
    }
    wave_lds_sync();
    if (c > 0) {
        w.ntest += c;
        const unsigned long long k = L.key[lane];
        if (any && k != NO_HIT_KEY) {
            // the walk ends occluded; its pending list no longer matters
            w.best = (int)(unsigned)(k & 0xffffffffull);
            w.best_t = w.tmax;
            return;
        }
        if (!any && k != NO_HIT_KEY) {
            w.best = (int)(unsigned)(k & 0xffffffffull);
            w.best_t = __uint_as_float((unsigned)(k >> 32));
        }
        const int ndf = L.ndefer[lane];
        if (ndf > 0) {
            w.nret += ndf;
            w.near = true;
        }
        if (L.npend[lane] > PENDING) w.deferred = true;
    }
}

// FP64 decision of the pending candidates pend[0..PENDING) (-1: none) of an
// FP32 walk whose nearest hit is (best, best_t): the inline re-test's
// acceptance and tie rules.  ANY: returns true as soon as one is a hit before
// tmax; otherwise returns true when one replaced the nearest hit.
template <int N, bool ANY>
__device__ __forceinline__ bool pending_resolve(const TreeArgs& A, const float* o, const float* d, const int* pend,
                                                float tmax, int* best, double* best_t)
{
    double o64[N], d64[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        o64[i] = o[i];
        d64[i] = d[i];
    }
    bool any = false;
    for (int k = 0; k < PENDING; ++k) {
        const int sid = pend[k];
        if (sid < 0) continue;
        double t64;
        if (!simplex_test64<N>(A.rec64 + (int64_t)sid * rec64_stride(N), o64, d64, &t64)) continue;
        if (ANY) {
            if ((float)t64 < tmax) return true;
            continue;
        }
        if (nearer<float>(t64, sid, *best_t, *best)) {
            *best_t = t64;
            *best = sid;
            any = true;
        }
    }
    return any;
}

// The whole walk of one ray in one call (first-hit queries and FP64 paths).
template <int N, typename T, bool ANY>
__device__ HitResult traverse(const TreeArgs& A, const T* o, const T* d, T tmax, int2* stack, const double* o64,
                              const double* d64)
{
    RayWalk<N, T, ANY> w;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        w.o[i] = o[i];
        w.d[i] = d[i];
    }
    if (!w.start(A, tmax)) return HitResult{-1, 0.0, false};
    while (!w.step(A, stack, o64, d64)) {
    }
    w.flush_counters(A);
    return w.result();
}

}  // namespace nsp
