// nsp_hull.hip -- convex hulls (2-6 dimensions) and Delaunay simplices (2-5
// dimensions) of float points with exact integer predicates; the conflict
// graph of the points left after the first insertions is computed on the GPU.
//
// Algorithm (compute_convex_hull / compute_delaunay, geometry/core/convex_hull.cpp):
// randomised incremental construction with conflict lists (de Berg, Cheong, van
// Kreveld, Overmars, Computational Geometry, 3rd ed., ch. 11, in M dimensions):
//  - the points are shuffled (std::shuffle, std::mt19937_64 seeded with the point
//    count), scaled to integers in [0, 2^30 - 1] (2^24 - 1 for Delaunay, whose
//    points are lifted to the paraboloid x_M = |x|^2) and duplicates dropped;
//  - an initial M-simplex from the first points that span the space;
//  - each point in turn: the facets it sees (strictly positive side) go, and
//    every horizon ridge (a ridge between a seen and an unseen facet) is joined
//    to the point.  A new facet's conflict points come from the two facets of its
//    ridge.
// The facets the construction makes do not depend on the order in which one
// step's seen facets or conflict points are visited, only on the point order.
// A point strictly inside the hull of the first K points can never see a later
// facet (later hulls contain this one).  So the host inserts the first K points
// with conflict lists over those points only, then k_hull_count / k_hull_fill
// test the points of the next stage, [K, 4K), against every facet of that hull
// on the GPU: points that see no facet are dropped, the others get their
// conflict lists, and the host inserts them; then the same for [4K, 16K), and
// so on.  The facets equal those of the plain construction for every input,
// degenerate ones included.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <memory>
#include <numeric>
#include <random>
#include <string>
#include <unordered_set>
#include <vector>

#include "nsp_exact.hpp"
#include "nsp_internal.hpp"

struct nsp_hull {
    int n = 0;             // dimension of the input points
    int per_facet = 0;     // vertices per output facet (n for hulls, n + 1 for Delaunay)
    int orthos_per = 0;    // unit normals per output facet (1 or n + 1)
    bool delaunay = false;
    std::vector<int32_t> vertices;  // [facets][per_facet], input point indices
    std::vector<double> orthos;     // [facets][orthos_per][n]
    std::vector<double> points;     // Delaunay: [count][n] integer coordinates used
    std::vector<int32_t> order;     // input index of each unique point, in insertion order
    int64_t stats[4] = {0, 0, 0, 0};
    double ms[3] = {0, 0, 0};
};

namespace nsp {
// from nsp_api.cpp
hipStream_t context_stream(nsp_context* ctx);
int context_device(nsp_context* ctx);
std::string* context_error(nsp_context* ctx);
}  // namespace nsp

namespace nsp {
namespace {

constexpr int ORDINARY_BITS = 30;   // convex_hull.cpp:96
constexpr int PARABOLOID_BITS = 24; // convex_hull.cpp:97
constexpr int64_t DEFAULT_FIRST = 4096;
constexpr int HULL_BLOCK = 256;

// Exact type of the predicates per dimension M: coordinates below 2^30 (or 2^24
// and their lifted squares below 5 * 2^48) keep every normal component and
// visibility value below 2^127 for M <= 4 and below 2^200 for M <= 6.
template <int M> struct ExactOf { using type = __int128; };
template <> struct ExactOf<5> { using type = I256; };
template <> struct ExactOf<6> { using type = I256; };

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// determinant of rows r0.. and the k listed columns of a row-major matrix (ld columns)
template <typename C>
C minor_det(const C* m, int ld, int r0, int k, const int* cols)
{
    if (k == 1) return m[r0 * ld + cols[0]];
    if (k == 2)
        return m[r0 * ld + cols[0]] * m[(r0 + 1) * ld + cols[1]] - m[r0 * ld + cols[1]] * m[(r0 + 1) * ld + cols[0]];
    C s = C(0);
    int sub[8];
    for (int j = 0; j < k; ++j) {
        int t = 0;
        for (int i = 0; i < k; ++i)
            if (i != j) sub[t++] = cols[i];
        const C term = m[r0 * ld + cols[j]] * minor_det(m, ld, r0 + 1, k - 1, sub);
        if (j & 1) s -= term;
        else s += term;
    }
    return s;
}

// A vector orthogonal to the M - 1 differences x[v_i] - x[v_0] (ortho_nn,
// geometry/core/linear_algebra.h): component j is the signed minor without column j.
template <int M, typename C>
void ortho_of(const int64_t* X, const int* v, C* out)
{
    C d[(M - 1) * M];
    for (int i = 1; i < M; ++i)
        for (int j = 0; j < M; ++j) d[(i - 1) * M + j] = C(X[(int64_t)v[i] * M + j] - X[(int64_t)v[0] * M + j]);
    int cols[8];
    for (int j = 0; j < M; ++j) {
        int t = 0;
        for (int i = 0; i < M; ++i)
            if (i != j) cols[t++] = i;
        const C m = minor_det(d, M, 0, M - 1, cols);
        out[j] = (j & 1) ? C(0) - m : m;
    }
}

template <int M, typename C>
__host__ __device__ inline C visible_value(const C* ortho, const int64_t* base, const int64_t* x)
{
    C s = ortho[0] * C(x[0] - base[0]);
#pragma unroll
    for (int j = 1; j < M; ++j) s += ortho[j] * C(x[j] - base[j]);
    return s;
}

// k independent differences among the first k rows (rank check by minors)
template <int M, typename C>
bool independent(const C* d, int k)
{
    int cols[8];
    // every k-subset of the M columns, as a bit mask
    for (unsigned mask = 0; mask < (1u << M); ++mask) {
        if (__builtin_popcount(mask) != k) continue;
        int t = 0;
        for (int j = 0; j < M; ++j)
            if (mask & (1u << j)) cols[t++] = j;
        if (sign_of(minor_det(d, M, 0, k, cols)) != 0) return true;
    }
    return false;
}

// ---------------------------------------------------------------------------
// GPU conflict pass

template <int M, typename C>
struct FacetDev {
    C ortho[M];
    int64_t base[M];
};

// One wave per point, its lanes strided over the facets: a hull of a few
// thousand points has tens of thousands of facets, so the points alone would
// not fill the GPU.
template <int M, typename C>
__global__ void __launch_bounds__(HULL_BLOCK) k_hull_count(const int64_t* __restrict__ X, const int32_t* __restrict__ pts,
                                                           int64_t npts, const FacetDev<M, C>* __restrict__ F,
                                                           int64_t nf, int32_t* __restrict__ count)
{
    const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = (int)(threadIdx.x & 63);
    if (w >= npts) return;
    int64_t x[M];
#pragma unroll
    for (int j = 0; j < M; ++j) x[j] = X[(int64_t)pts[w] * M + j];
    int c = 0;
    for (int64_t f = lane; f < nf; f += 64) c += sign_of(visible_value<M, C>(F[f].ortho, F[f].base, x)) > 0;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) c += __shfl_xor(c, m, 64);
    if (lane == 0) count[w] = c;
}

// The same tests again, each wave writing its point's facets in facet order
// (ballot and prefix count per 64 facets).
template <int M, typename C>
__global__ void __launch_bounds__(HULL_BLOCK) k_hull_fill(const int64_t* __restrict__ X, const int32_t* __restrict__ pts,
                                                          int64_t npts, const FacetDev<M, C>* __restrict__ F, int64_t nf,
                                                          const int64_t* __restrict__ offset,
                                                          int32_t* __restrict__ lists)
{
    const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = (int)(threadIdx.x & 63);
    if (w >= npts) return;
    int64_t o = offset[w];
    const int64_t o1 = offset[w + 1];
    if (o == o1) return;
    int64_t x[M];
#pragma unroll
    for (int j = 0; j < M; ++j) x[j] = X[(int64_t)pts[w] * M + j];
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    for (int64_t f0 = 0; f0 < nf && o < o1; f0 += 64) {
        const int64_t f = f0 + lane;
        const bool vis = f < nf && sign_of(visible_value<M, C>(F[f].ortho, F[f].base, x)) > 0;
        const unsigned long long m = __ballot(vis);
        if (vis) lists[o + __popcll(m & below)] = (int32_t)f;
        o += __popcll(m);
    }
}

// ---------------------------------------------------------------------------
// incremental construction

template <int M>
struct Builder {
    using C = typename ExactOf<M>::type;
    struct Facet {
        int v[M];
        int link[M];  // link[r]: the facet across the ridge without v[r]
        C ortho[M];   // outward
        double od[M]; // ortho rounded to double (the filter of visible_sign)
        std::vector<int> conf;
        bool alive = true;
        bool seen = false;
    };

    const std::vector<int64_t>& X;  // [P][M]
    int P;
    std::vector<Facet> F;
    std::vector<std::vector<int>> pconf;  // facets a point sees (dead ones skipped)
    std::vector<uint8_t> enabled;
    std::vector<int> mark;
    int stamp = 0;
    int64_t alive_count = 0;
    // device buffers of the conflict passes, kept across the stages (DevBuf::alloc
    // keeps a buffer that is large enough); the points go up once
    DevBuf dX, dP, dF, dC, dO, dL;
    bool x_uploaded = false;

    Builder(const std::vector<int64_t>& x, int p) : X(x), P(p), pconf(p), enabled(p, 1), mark(p, 0) {}

    C visible(const Facet& f, int q) const
    {
        return visible_value<M, C>(f.ortho, &X[(int64_t)f.v[0] * M], &X[(int64_t)q * M]);
    }

    // Sign of visible(f, q).  The double dot product of the rounded normal with
    // the (exactly converted) difference is within (M + 3) 2^-53 sum|terms| of
    // the exact value (normal rounding 3 ulp at most, products 1, sums M - 1), so
    // a value beyond 16 2^-53 sum|terms| has the exact sign; otherwise the
    // integer test decides.
    int visible_sign(const Facet& f, int q) const
    {
        const int64_t* b = &X[(int64_t)f.v[0] * M];
        const int64_t* x = &X[(int64_t)q * M];
        double s = 0, a = 0;
        for (int j = 0; j < M; ++j) {
            const double t = f.od[j] * (double)(x[j] - b[j]);
            s += t;
            a += std::fabs(t);
        }
        if (std::fabs(s) > a * (16.0 / 9007199254740992.0)) return s > 0 ? 1 : -1;
        return sign_of(visible(f, q));
    }

    // the facet's normal, pointing away from `inner` (FacetInteger, geometry/core/facet.h:178-207)
    void orient(Facet& f, int inner, const Facet* horizon)
    {
        ortho_of<M, C>(X.data(), f.v, f.ortho);
        const int s = sign_of(visible(f, inner));
        bool flip = s > 0;
        if (s == 0) {
            if (!horizon) raise(NSP_E_INVALID, "convex hull: degenerate initial simplex");
            // coplanar with the horizon facet: the same direction
            for (int j = 0; j < M; ++j) {
                const int a = sign_of(f.ortho[j]), b = sign_of(horizon->ortho[j]);
                if ((a > 0 && b < 0) || (a < 0 && b > 0)) flip = true;
            }
        }
        if (flip)
            for (int j = 0; j < M; ++j) f.ortho[j] = C(0) - f.ortho[j];
        for (int j = 0; j < M; ++j) f.od[j] = to_double(f.ortho[j]);
    }

    // pair up the ridges of the listed facets (all ridges, or those without `apex`)
    void connect(const std::vector<int>& ids, int apex)
    {
        struct R {
            std::array<int, M> key;
            int f, r;
        };
        std::vector<R> rs;
        for (int id : ids) {
            const Facet& f = F[id];
            for (int r = 0; r < M; ++r) {
                if (f.v[r] == apex) continue;
                R e;
                int t = 0;
                for (int i = 0; i < M; ++i)
                    if (i != r) e.key[t++] = f.v[i];
                std::sort(e.key.begin(), e.key.begin() + (M - 1));
                e.key[M - 1] = 0;
                e.f = id;
                e.r = r;
                rs.push_back(e);
            }
        }
        std::sort(rs.begin(), rs.end(), [](const R& a, const R& b) { return a.key < b.key; });
        for (size_t i = 0; i < rs.size(); i += 2) {
            if (i + 1 >= rs.size() || rs[i].key != rs[i + 1].key) raise(NSP_E_INVALID, "convex hull: open ridge");
            if (i + 2 < rs.size() && rs[i + 2].key == rs[i].key) raise(NSP_E_INVALID, "convex hull: ridge of three facets");
            F[rs[i].f].link[rs[i].r] = rs[i + 1].f;
            F[rs[i + 1].f].link[rs[i + 1].r] = rs[i].f;
        }
    }

    // find_simplex_points (convex_hull.cpp:255-298): point 0, then for each next
    // vertex the first later point whose difference keeps the set independent
    std::array<int, M + 1> initial_simplex() const
    {
        std::array<int, M + 1> s;
        s[0] = 0;
        std::vector<C> d((size_t)M * M);
        int next = 1;
        for (int k = 1; k <= M; ++k) {
            int p = next;
            for (; p < P; ++p) {
                for (int j = 0; j < M; ++j) d[(k - 1) * M + j] = C(X[(int64_t)p * M + j] - X[j]);
                if (independent<M, C>(d.data(), k)) break;
            }
            if (p >= P)
                raise(NSP_E_INVALID, "point " + std::to_string(k + 1) + " of " + std::to_string(M) +
                                         "-simplex not found");
            s[k] = p;
            next = p + 1;
        }
        return s;
    }

    void start(int first)
    {
        if (P < M + 1)
            raise(NSP_E_INVALID, "Error point count " + std::to_string(P) + " for convex hull in " +
                                     std::to_string(M) + " dimensions");
        const std::array<int, M + 1> s = initial_simplex();
        std::vector<int> ids;
        for (int f = 0; f <= M; ++f) {
            Facet nf;
            int t = 0;
            for (int i = 0; i <= M; ++i)
                if (i != f) nf.v[t++] = s[i];
            orient(nf, s[f], nullptr);
            F.push_back(std::move(nf));
            ids.push_back((int)F.size() - 1);
        }
        connect(ids, -1);
        alive_count = M + 1;
        for (int v : s) enabled[v] = 0;
        for (int id : ids)
            for (int q = 0; q < first; ++q)
                if (enabled[q] && visible_sign(F[id], q) > 0) {
                    F[id].conf.push_back(q);
                    pconf[q].push_back(id);
                }
    }

    void insert(int p)
    {
        std::vector<int> seen;
        for (int id : pconf[p])
            if (F[id].alive) seen.push_back(id);
        pconf[p].clear();
        pconf[p].shrink_to_fit();
        if (seen.empty()) return;  // inside
        if ((int64_t)seen.size() >= alive_count) raise(NSP_E_INVALID, "All facets are visible from the point");
        for (int id : seen) F[id].seen = true;
        std::vector<int> created;
        for (int id : seen) {
            for (int r = 0; r < M; ++r) {
                const int g = F[id].link[r];
                if (F[g].seen) continue;
                int gi = 0;
                while (F[g].link[gi] != id) ++gi;
                Facet nf;
                for (int i = 0; i < M; ++i) nf.v[i] = F[id].v[i];
                nf.v[r] = p;
                orient(nf, F[g].v[gi], &F[g]);
                // conflict points of the new facet: those of the two facets of its ridge
                ++stamp;
                for (int q : F[id].conf) {
                    mark[q] = stamp;
                    if (q != p && visible_sign(nf, q) > 0) nf.conf.push_back(q);
                }
                for (int q : F[g].conf)
                    if (mark[q] != stamp && q != p && visible_sign(nf, q) > 0) nf.conf.push_back(q);
                F.push_back(std::move(nf));
                const int nid = (int)F.size() - 1;
                F[nid].link[r] = g;
                F[g].link[gi] = nid;
                created.push_back(nid);
            }
        }
        for (int id : seen) {
            F[id].alive = false;
            std::vector<int>().swap(F[id].conf);
        }
        alive_count += (int64_t)created.size() - (int64_t)seen.size();
        connect(created, p);
        for (int nid : created)
            for (int q : F[nid].conf) pconf[q].push_back(nid);
    }

    // conflict graph of points [lo, hi) against the current hull, on the GPU
    void gpu_conflicts(int lo, int hi, hipStream_t s, int64_t* dropped)
    {
        std::vector<int32_t> pts;
        for (int q = lo; q < hi; ++q)
            if (enabled[q]) pts.push_back(q);
        if (pts.empty()) return;
        std::vector<int> ids;
        std::vector<FacetDev<M, C>> fd;
        for (int id = 0; id < (int)F.size(); ++id) {
            if (!F[id].alive) continue;
            FacetDev<M, C> d;
            for (int j = 0; j < M; ++j) {
                d.ortho[j] = F[id].ortho[j];
                d.base[j] = X[(int64_t)F[id].v[0] * M + j];
            }
            fd.push_back(d);
            ids.push_back(id);
        }
        const int64_t np = (int64_t)pts.size(), nf = (int64_t)fd.size();
        dX.alloc(sizeof(int64_t) * X.size());
        dP.alloc(sizeof(int32_t) * np);
        dF.alloc(sizeof(FacetDev<M, C>) * nf);
        dC.alloc(sizeof(int32_t) * np);
        dO.alloc(sizeof(int64_t) * (np + 1));
        if (!x_uploaded) {
            NSP_HIP(hipMemcpyAsync(dX.p, X.data(), sizeof(int64_t) * X.size(), hipMemcpyHostToDevice, s));
            x_uploaded = true;
        }
        NSP_HIP(hipMemcpyAsync(dP.p, pts.data(), sizeof(int32_t) * np, hipMemcpyHostToDevice, s));
        NSP_HIP(hipMemcpyAsync(dF.p, fd.data(), sizeof(FacetDev<M, C>) * nf, hipMemcpyHostToDevice, s));
        const unsigned grid = (unsigned)((np * 64 + HULL_BLOCK - 1) / HULL_BLOCK);
        hipLaunchKernelGGL((k_hull_count<M, C>), dim3(grid), dim3(HULL_BLOCK), 0, s, dX.as<int64_t>(),
                           dP.as<int32_t>(), np, dF.as<FacetDev<M, C>>(), nf, dC.as<int32_t>());
        NSP_HIP(hipGetLastError());
        std::vector<int32_t> cnt(np);
        NSP_HIP(hipMemcpyAsync(cnt.data(), dC.p, sizeof(int32_t) * np, hipMemcpyDeviceToHost, s));
        NSP_HIP(hipStreamSynchronize(s));
        std::vector<int64_t> off(np + 1, 0);
        for (int64_t i = 0; i < np; ++i) off[i + 1] = off[i] + cnt[i];
        const int64_t total = off[np];
        std::vector<int32_t> lists(total);
        if (total > 0) {
            dL.alloc(sizeof(int32_t) * total);
            NSP_HIP(hipMemcpyAsync(dO.p, off.data(), sizeof(int64_t) * (np + 1), hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL((k_hull_fill<M, C>), dim3(grid), dim3(HULL_BLOCK), 0, s, dX.as<int64_t>(),
                               dP.as<int32_t>(), np, dF.as<FacetDev<M, C>>(), nf, dO.as<int64_t>(),
                               dL.as<int32_t>());
            NSP_HIP(hipGetLastError());
            NSP_HIP(hipMemcpyAsync(lists.data(), dL.p, sizeof(int32_t) * total, hipMemcpyDeviceToHost, s));
            NSP_HIP(hipStreamSynchronize(s));
        }
        for (int64_t i = 0; i < np; ++i) {
            const int q = pts[i];
            if (cnt[i] == 0) {
                enabled[q] = 0;
                ++*dropped;
                continue;
            }
            for (int64_t o = off[i]; o < off[i + 1]; ++o) {
                const int id = ids[lists[o]];
                pconf[q].push_back(id);
                F[id].conf.push_back(q);
            }
        }
    }

    // Stages [first, 4 first), [4 first, 16 first), ...: at each stage's start the
    // GPU gives the conflicts of that stage's points only, so the host never
    // tracks points of later stages, and each later pass meets a larger hull that
    // more points lie inside.
    void run(int first, hipStream_t s, nsp_hull* out)
    {
        double t0 = now_ms();
        first = std::max(M + 1, std::min(first, P));
        start(first);
        for (int q = 0; q < first; ++q)
            if (enabled[q]) insert(q);
        out->ms[0] = now_ms() - t0;
        for (int lo = first; lo < P;) {
            const int hi = (int)std::min<int64_t>(P, (int64_t)lo * 4);
            double t1 = now_ms();
            gpu_conflicts(lo, hi, s, &out->stats[1]);
            double t2 = now_ms();
            for (int q = lo; q < hi; ++q)
                if (enabled[q]) insert(q);
            out->ms[1] += t2 - t1;
            out->ms[2] += now_ms() - t2;
            lo = hi;
        }
        out->stats[2] = (int64_t)F.size();
    }
};

// shuffle_and_convert_to_unique_integer (convex_hull.cpp:661-711)
template <int N>
void to_unique_integers(const float* src, int64_t count, int64_t max_value, std::vector<std::array<int64_t, N>>* pts,
                        std::vector<int32_t>* order)
{
    std::vector<int32_t> perm((size_t)count);
    std::iota(perm.begin(), perm.end(), 0);
    std::shuffle(perm.begin(), perm.end(), std::mt19937_64((uint64_t)count));
    float mn[N], mx[N];
    for (int j = 0; j < N; ++j) mn[j] = mx[j] = src[j];
    for (int64_t i = 1; i < count; ++i)
        for (int j = 0; j < N; ++j) {
            mn[j] = std::min(mn[j], src[i * N + j]);
            mx[j] = std::max(mx[j], src[i * N + j]);
        }
    float max_f = mx[0] - mn[0];
    for (int j = 1; j < N; ++j) max_f = std::max(max_f, mx[j] - mn[j]);
    const double max_d = max_f;
    if (max_d == 0) raise(NSP_E_INVALID, "all points are equal to each other");
    const double scale = (double)max_value / max_d;
    struct Hash {
        size_t operator()(const std::array<int64_t, N>& a) const
        {
            uint64_t h = 1469598103934665603ull;
            for (int64_t v : a) h = (h ^ (uint64_t)v) * 1099511628211ull;
            return (size_t)h;
        }
    };
    std::unordered_set<std::array<int64_t, N>, Hash> seen((size_t)count);
    pts->clear();
    order->clear();
    for (int64_t i = 0; i < count; ++i) {
        const int32_t r = perm[i];
        std::array<int64_t, N> v;
        for (int j = 0; j < N; ++j) {
            const float diff = src[(int64_t)r * N + j] - mn[j];
            const long long ll = std::llround((double)diff * scale);
            if (ll < 0 || ll > max_value) raise(NSP_E_INVALID, "points preprocessing error: value < 0 || value > MAX");
            v[j] = ll;
        }
        if (seen.insert(v).second) {
            pts->push_back(v);
            order->push_back(r);
        }
    }
}

template <int N>
void convex_hull_t(const float* src, int64_t count, int64_t first, hipStream_t s, nsp_hull* out)
{
    std::vector<std::array<int64_t, N>> pts;
    to_unique_integers<N>(src, count, (int64_t(1) << ORDINARY_BITS) - 1, &pts, &out->order);
    const int P = (int)pts.size();
    std::vector<int64_t> X((size_t)P * N);
    for (int i = 0; i < P; ++i)
        for (int j = 0; j < N; ++j) X[(size_t)i * N + j] = pts[i][j];
    Builder<N> b(X, P);
    b.run((int)first, s, out);
    out->stats[0] = P;
    for (const auto& f : b.F) {
        if (!f.alive) continue;
        double o[N], len = 0;
        for (int j = 0; j < N; ++j) {
            o[j] = to_double(f.ortho[j]);
            len += o[j] * o[j];
        }
        len = std::sqrt(len);
        for (int j = 0; j < N; ++j) {
            out->vertices.push_back(out->order[f.v[j]]);
            out->orthos.push_back(o[j] / len);
        }
    }
}

template <int N>
void delaunay_t(const float* src, int64_t count, int64_t first, hipStream_t s, nsp_hull* out)
{
    constexpr int M = N + 1;
    std::vector<std::array<int64_t, N>> pts;
    to_unique_integers<N>(src, count, (int64_t(1) << PARABOLOID_BITS) - 1, &pts, &out->order);
    const int P = (int)pts.size();
    // the paraboloid x_M = |x|^2 (paraboloid_convex_hull, convex_hull.cpp:731-744)
    std::vector<int64_t> X((size_t)P * M);
    for (int i = 0; i < P; ++i) {
        int64_t sq = 0;
        for (int j = 0; j < N; ++j) {
            X[(size_t)i * M + j] = pts[i][j];
            sq += pts[i][j] * pts[i][j];
        }
        X[(size_t)i * M + N] = sq;
    }
    Builder<M> b(X, P);
    b.run((int)first, s, out);
    out->stats[0] = P;
    std::vector<int64_t> Xd((size_t)P * N);
    for (int i = 0; i < P; ++i)
        for (int j = 0; j < N; ++j) Xd[(size_t)i * N + j] = pts[i][j];
    using CD = typename ExactOf<N>::type;
    for (const auto& f : b.F) {
        // lower facets only: the last normal coordinate negative
        if (!f.alive || sign_of(f.ortho[N]) >= 0) continue;
        for (int j = 0; j < M; ++j) out->vertices.push_back(out->order[f.v[j]]);
        // per vertex r, the unit normal of the face without r, away from r
        for (int r = 0; r < M; ++r) {
            int v[N];
            int t = 0;
            for (int i = 0; i < M; ++i)
                if (i != r) v[t++] = f.v[i];
            CD o[N];
            ortho_of<N, CD>(Xd.data(), v, o);
            const int sg = sign_of(visible_value<N, CD>(o, &Xd[(size_t)v[0] * N], &Xd[(size_t)f.v[r] * N]));
            if (sg == 0) raise(NSP_E_INVALID, "Delaunay: degenerate simplex");
            double od[N], len = 0;
            for (int j = 0; j < N; ++j) {
                od[j] = to_double(o[j]) * (sg > 0 ? -1.0 : 1.0);
                len += od[j] * od[j];
            }
            len = std::sqrt(len);
            for (int j = 0; j < N; ++j) out->orthos.push_back(od[j] / len);
        }
    }
    out->points.assign((size_t)count * N, 0.0);
    for (int i = 0; i < P; ++i)
        for (int j = 0; j < N; ++j) out->points[(size_t)out->order[i] * N + j] = (double)pts[i][j];
}

int hull_entry(nsp_context* ctx, int n, const float* points, int64_t count, int64_t first, bool delaunay,
               nsp_hull** out)
{
    if (!ctx || !out) return NSP_E_INVALID;
    *out = nullptr;
    std::string* err = context_error(ctx);
    try {
        NSP_HIP(hipSetDevice(context_device(ctx)));
        if (!points && count > 0) raise(NSP_E_INVALID, "points is NULL");
        if (count <= 0) raise(NSP_E_INVALID, "no points for convex hull");
        if (count > INT32_MAX) raise(NSP_E_INVALID, "too many points");
        if (first < 0) raise(NSP_E_INVALID, "first_points must be >= 0");
        if (delaunay ? (n < 2 || n > 5) : (n < 2 || n > 6))
            raise(NSP_E_UNSUPPORTED, delaunay ? "Delaunay dimension must be in [2, 5]"
                                              : "convex hull dimension must be in [2, 6]");
        for (int64_t i = 0; i < count * n; ++i)
            if (!std::isfinite(points[i])) raise(NSP_E_INVALID, "point coordinates must be finite");
        auto h = std::make_unique<nsp_hull>();
        h->n = n;
        h->delaunay = delaunay;
        h->per_facet = delaunay ? n + 1 : n;
        h->orthos_per = delaunay ? n + 1 : 1;
        const int64_t f = first > 0 ? first : DEFAULT_FIRST;
        hipStream_t s = context_stream(ctx);
        if (delaunay) {
            switch (n) {
            case 2: delaunay_t<2>(points, count, f, s, h.get()); break;
            case 3: delaunay_t<3>(points, count, f, s, h.get()); break;
            case 4: delaunay_t<4>(points, count, f, s, h.get()); break;
            case 5: delaunay_t<5>(points, count, f, s, h.get()); break;
            }
        } else {
            switch (n) {
            case 2: convex_hull_t<2>(points, count, f, s, h.get()); break;
            case 3: convex_hull_t<3>(points, count, f, s, h.get()); break;
            case 4: convex_hull_t<4>(points, count, f, s, h.get()); break;
            case 5: convex_hull_t<5>(points, count, f, s, h.get()); break;
            case 6: convex_hull_t<6>(points, count, f, s, h.get()); break;
            }
        }
        *out = h.release();
        return NSP_OK;
    } catch (const Error& e) {
        *err = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        *err = "host allocation failed";
        return NSP_E_NOMEM;
    }
}

}  // namespace
}  // namespace nsp

using namespace nsp;

int nsp_convex_hull(nsp_context* ctx, int n, const float* points, int64_t count, int64_t first_points, nsp_hull** out)
{
    return hull_entry(ctx, n, points, count, first_points, false, out);
}

int nsp_delaunay(nsp_context* ctx, int n, const float* points, int64_t count, int64_t first_points, nsp_hull** out)
{
    return hull_entry(ctx, n, points, count, first_points, true, out);
}

int nsp_hull_info(const nsp_hull* h, int* n, int64_t* facets, int* vertices_per_facet, int* orthos_per_facet,
                  int64_t* unique_points)
{
    if (!h) return NSP_E_INVALID;
    if (n) *n = h->n;
    if (facets) *facets = (int64_t)(h->vertices.size() / h->per_facet);
    if (vertices_per_facet) *vertices_per_facet = h->per_facet;
    if (orthos_per_facet) *orthos_per_facet = h->orthos_per;
    if (unique_points) *unique_points = (int64_t)h->order.size();
    return NSP_OK;
}

int nsp_hull_copy(const nsp_hull* h, int32_t* vertices, double* orthos, double* points, int32_t* order)
{
    if (!h) return NSP_E_INVALID;
    if (points && !h->delaunay) return NSP_E_INVALID;
    if (vertices) std::copy(h->vertices.begin(), h->vertices.end(), vertices);
    if (orthos) std::copy(h->orthos.begin(), h->orthos.end(), orthos);
    if (points) std::copy(h->points.begin(), h->points.end(), points);
    if (order) std::copy(h->order.begin(), h->order.end(), order);
    return NSP_OK;
}

int nsp_hull_stats(const nsp_hull* h, int64_t* dropped, int64_t* facets_created, double* ms)
{
    if (!h) return NSP_E_INVALID;
    if (dropped) *dropped = h->stats[1];
    if (facets_created) *facets_created = h->stats[2];
    if (ms)
        for (int i = 0; i < 3; ++i) ms[i] = h->ms[i];
    return NSP_OK;
}

void nsp_hull_destroy(nsp_hull* h) { delete h; }
