// The reconstruction test of the reference (geometry/test/test_reconstruction.cpp:
// 184-281) through the C++ drop-in (painter/reconstruction.h): Cocone and
// BoundCocone (rho 0.3, alpha 0.14) of a sphere with a notch, then of it with
// its 2^n clones shifted by 3, in 2 and 3 dimensions, with the facet counts the
// reference requires; BoundCocone of the bounded notched sphere.  The points
// follow generate_points_sphere_with_notch (geometry/objects/points.cpp:
// 131-236): std::mt19937_64 seeded with the point count, uniform points of the
// cube [-1, 1]^n inside the unit ball projected on the sphere, the last
// coordinate z > 0 scaled by 1 - |z^5| / 2, points equal after rounding at
// 1e-5 dropped.  Built with the host compiler; prints "OK" and the counts.
#include <nsp.h>

#include "painter/reconstruction.h"

#include <cmath>
#include <cstdio>
#include <map>
#include <random>
#include <set>
#include <string>
#include <vector>

namespace {

template <size_t N>
std::vector<Vector<N, float>> sphere_with_notch(unsigned count, bool bound)
{
    std::mt19937_64 engine(count);
    std::uniform_real_distribution<double> urd(-1.0, 1.0);
    std::set<std::array<long, N>> seen;
    std::vector<Vector<N, float>> out;
    while (out.size() < count) {
        double v[N];
        double len2;
        do {
            do {
                len2 = 0;
                for (size_t i = 0; i < N; ++i) {
                    v[i] = urd(engine);
                    len2 += v[i] * v[i];
                }
            } while (len2 > 1);
            const double len = std::sqrt(len2);
            for (size_t i = 0; i < N; ++i) v[i] /= len;
        } while (bound && v[N - 1] < -0.3);
        const double z = v[N - 1];
        if (z > 0) v[N - 1] *= 1 - std::abs(0.5 * std::pow(z, 5));
        std::array<long, N> key;
        for (size_t i = 0; i < N; ++i) key[i] = std::lround(v[i] * 100000);
        if (!seen.insert(key).second) continue;
        Vector<N, float> p;
        for (size_t i = 0; i < N; ++i) p[i] = (float)v[i];
        out.push_back(p);
    }
    return out;
}

template <size_t N>
std::vector<Vector<N, float>> clones(const std::vector<Vector<N, float>>& points)
{
    std::vector<Vector<N, float>> out(points);
    for (unsigned k = 0; k < (1u << N); ++k)
        for (const auto& p : points) {
            Vector<N, float> q;
            for (size_t i = 0; i < N; ++i) q[i] = p[i] + (((k >> i) & 1) ? 3.0f : -3.0f);
            out.push_back(q);
        }
    return out;
}

bool ridges_twice(const std::vector<std::array<int, 3>>& facets)
{
    std::map<std::array<int, 2>, int> cnt;
    for (const auto& f : facets)
        for (int r = 0; r < 3; ++r) {
            std::array<int, 2> e;
            int c = 0;
            for (int q = 0; q < 3; ++q)
                if (q != r) e[c++] = f[q];
            ++cnt[e];
        }
    for (const auto& kv : cnt)
        if (kv.second != 2) return false;
    return true;
}

template <size_t N>
bool check(const char* name, const std::vector<Vector<N, float>>& points, bool cocone, unsigned lo, unsigned hi,
           unsigned blo, unsigned bhi)
{
    ProgressRatio progress(nullptr);
    std::unique_ptr<ManifoldConstructor<N>> sr = create_manifold_constructor(points, &progress);
    if (cocone) {
        std::vector<vec<N>> normals(10000);
        std::vector<std::array<int, N>> facets(10000);
        sr->cocone(&normals, &facets, &progress);
        std::printf("%s: %zu points, Cocone %zu facets, expected [%u, %u]\n", name, points.size(), facets.size(), lo,
                    hi);
        if (!(lo <= facets.size() && facets.size() <= hi) || normals.size() != points.size()) return false;
        if constexpr (N == 3) {
            if (!ridges_twice(facets)) return false;
        }
    }
    std::vector<vec<N>> normals(10000);
    std::vector<std::array<int, N>> facets(10000);
    sr->bound_cocone(0.3, 0.14, &normals, &facets, &progress);
    std::printf("%s: %zu points, BoundCocone %zu facets, expected [%u, %u]\n", name, points.size(), facets.size(), blo,
                bhi);
    return blo <= facets.size() && facets.size() <= bhi && normals.size() == points.size();
}

template <size_t N>
bool all_tests(unsigned count)
{
    const unsigned lo = N == 2 ? count : 2 * count - 4;
    const unsigned hi = lo;
    const unsigned blo = std::lround(0.9 * lo), bhi = std::lround(1.1 * hi);
    const unsigned objects = 1 + (1u << N);
    const std::string space = std::to_string(N) + "-space";
    for (int bound = 0; bound < 2; ++bound) {
        const auto p = sphere_with_notch<N>(count, bound);
        const std::string name = space + (bound ? ", bounded" : ", unbounded");
        if (!check<N>((name + ", 1 object").c_str(), p, !bound, lo, hi, blo, bhi)) return false;
        if (!check<N>((name + ", " + std::to_string(objects) + " objects").c_str(), clones<N>(p), !bound,
                      lo * objects, hi * objects, blo * objects, bhi * objects))
            return false;
    }
    return true;
}

}  // namespace

int main()
{
    try {
        if (!all_tests<2>(600) || !all_tests<3>(2500)) {
            std::printf("FAIL counts\n");
            return 1;
        }
        // a constructor for Cocone only refuses BoundCocone
        ProgressRatio progress(nullptr);
        auto c = create_manifold_constructor_cocone<3>(sphere_with_notch<3>(2000, false), &progress);
        std::vector<vec<3>> normals;
        std::vector<std::array<int, 3>> facets;
        c->cocone(&normals, &facets, &progress);
        if (facets.size() != 2 * 2000 - 4 || c->delaunay_objects().empty()) {
            std::printf("FAIL cocone-only constructor\n");
            return 1;
        }
        try {
            dynamic_cast<ManifoldConstructor<3>&>(*c).bound_cocone(0.3, 0.14, &normals, &facets, &progress);
            std::printf("FAIL no error\n");
            return 1;
        } catch (const ErrorException& e) {
            if (std::string(e.what()) != "Manifold constructor created for Cocone and not for BoundCocone") {
                std::printf("FAIL message: %s\n", e.what());
                return 1;
            }
        }
    } catch (const std::exception& e) {
        std::printf("FAIL %s\n", e.what());
        return 1;
    }
    std::printf("OK\n");
    return 0;
}
