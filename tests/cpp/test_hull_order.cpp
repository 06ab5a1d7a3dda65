// The insertion order nsp_convex_hull reports is the reference's
// (shuffle_and_convert_to_unique_integer, geometry/core/convex_hull.cpp:661-711):
// std::shuffle of 0..count-1 with std::mt19937_64(count), keeping the first
// point of each group of equal points.  The C++ drop-in (painter/convex_hull.h)
// gives the same facets.  Built with the host compiler against include/nsp.h;
// prints "OK <unique> <facets> <delaunay simplices>".
#include <nsp.h>

#include "painter/convex_hull.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>
#include <random>
#include <vector>

int main()
{
    nsp_context* ctx = nullptr;
    if (nsp_context_create(0, &ctx) != NSP_OK) {
        std::printf("FAIL context\n");
        return 1;
    }
    // points on a 3-dimensional grid, with points 7, 9 and 30 repeated later
    const int count = 400, n = 3;
    std::vector<float> p((size_t)count * n);
    for (int i = 0; i < count; ++i)
        for (int j = 0; j < n; ++j) p[(size_t)i * n + j] = (float)((i * (j + 3) * 7919 + j * 104729) % 1009);
    const int copies[][2] = {{100, 7}, {200, 7}, {150, 9}, {399, 30}};
    for (auto& c : copies)
        for (int j = 0; j < n; ++j) p[(size_t)c[0] * n + j] = p[(size_t)c[1] * n + j];
    std::vector<int> dup_of(count, -1);
    for (auto& c : copies) dup_of[c[0]] = c[1];

    nsp_hull* h = nullptr;
    if (nsp_convex_hull(ctx, n, p.data(), count, 64, &h) != NSP_OK) {
        std::printf("FAIL hull: %s\n", nsp_last_error(ctx));
        return 1;
    }
    int64_t facets = 0, unique = 0;
    nsp_hull_info(h, nullptr, &facets, nullptr, nullptr, &unique);
    std::vector<int32_t> order((size_t)unique);
    nsp_hull_copy(h, nullptr, nullptr, nullptr, order.data());

    std::vector<int> perm(count);
    std::iota(perm.begin(), perm.end(), 0);
    std::shuffle(perm.begin(), perm.end(), std::mt19937_64(count));
    // the first of each group in shuffled order survives
    std::vector<int> group(count);
    for (int i = 0; i < count; ++i) group[i] = dup_of[i] >= 0 ? dup_of[i] : i;
    std::vector<char> taken(count, 0);
    std::vector<int> expect;
    for (int r : perm)
        if (!taken[group[r]]) {
            taken[group[r]] = 1;
            expect.push_back(r);
        }
    if (expect.size() != order.size()) {
        std::printf("FAIL unique %zu vs %zu\n", order.size(), expect.size());
        return 1;
    }
    for (size_t i = 0; i < expect.size(); ++i)
        if (expect[i] != order[i]) {
            std::printf("FAIL order at %zu: %d vs %d\n", i, order[i], expect[i]);
            return 1;
        }
    std::vector<int32_t> verts((size_t)facets * n);
    nsp_hull_copy(h, verts.data(), nullptr, nullptr, nullptr);
    nsp_hull_destroy(h);
    nsp_context_destroy(ctx);

    // the drop-in: compute_convex_hull / compute_delaunay on the same points
    std::vector<Vector<3, float>> pts(count);
    for (int i = 0; i < count; ++i)
        for (int j = 0; j < n; ++j) pts[i][j] = p[(size_t)i * n + j];
    std::vector<ConvexHullFacet<3>> ch;
    ProgressRatio progress(nullptr);
    compute_convex_hull(pts, &ch, &progress);
    if ((int64_t)ch.size() != facets) {
        std::printf("FAIL drop-in facets %zu vs %lld\n", ch.size(), (long long)facets);
        return 1;
    }
    for (size_t f = 0; f < ch.size(); ++f) {
        double len = 0;
        for (int j = 0; j < n; ++j) {
            if (ch[f].vertices()[j] != verts[f * n + j]) {
                std::printf("FAIL drop-in facet %zu\n", f);
                return 1;
            }
            len += ch[f].ortho()[j] * ch[f].ortho()[j];
        }
        if (std::abs(len - 1) > 1e-12) {
            std::printf("FAIL drop-in normal %zu\n", f);
            return 1;
        }
    }
    std::vector<vec<3>> dpts;
    std::vector<DelaunaySimplex<3>> ds;
    compute_delaunay(pts, &dpts, &ds, &progress);
    if (ds.empty() || dpts.size() != pts.size()) {
        std::printf("FAIL drop-in delaunay\n");
        return 1;
    }
    bool threw = false;
    try {
        std::vector<Vector<3, float>> same(10, Vector<3, float>{1, 1, 1});
        compute_convex_hull(same, &ch, &progress);
    } catch (const ErrorException&) {
        threw = true;
    }
    if (!threw) {
        std::printf("FAIL drop-in error\n");
        return 1;
    }
    std::printf("OK %lld %lld %zu\n", (long long)unique, (long long)facets, ds.size());
    return 0;
}
