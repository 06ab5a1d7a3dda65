// convex_hull.h -- C++17 drop-in for Math Viewer's convex hull and Delaunay
// entry points (src/geometry/core/convex_hull.h:26-74) on libnsp: the same
// class and function names and argument meaning, errors thrown as
// ErrorException with the reference's messages.  The hull is built with exact
// integer predicates; after the first 4096 points the GPU of the painter's
// context computes the conflicts of the remaining points in one pass.
#pragma once

#include <array>
#include <cstdint>
#include <vector>

#include "painter.h"

template <size_t N>
using vec = Vector<N, double>;

// ConvexHullFacet<N> (convex_hull.h:26-44): n vertex indices and the unit outward normal.
template <size_t N>
class ConvexHullFacet {
    std::array<int, N> m_indices;
    vec<N> m_ortho;

public:
    ConvexHullFacet(const std::array<int, N>& indices, const vec<N>& ortho) : m_indices(indices), m_ortho(ortho) {}
    const std::array<int, N>& vertices() const { return m_indices; }
    const vec<N>& ortho() const { return m_ortho; }
};

// DelaunaySimplex<N> (convex_hull.h:46-65): n + 1 vertex indices; ortho(i) is the
// unit normal of the face without vertex i, pointing away from it.
template <size_t N>
class DelaunaySimplex {
    std::array<int, N + 1> m_indices;
    std::array<vec<N>, N + 1> m_orthos;

public:
    DelaunaySimplex(const std::array<int, N + 1>& indices, const std::array<vec<N>, N + 1>& orthos)
        : m_indices(indices), m_orthos(orthos)
    {
    }
    const std::array<int, N + 1>& vertices() const { return m_indices; }
    const vec<N>& ortho(unsigned i) const { return m_orthos[i]; }
};

namespace nsp_painter {

struct HullHandle {
    nsp_hull* h = nullptr;
    ~HullHandle() { nsp_hull_destroy(h); }
};

template <size_t N>
void hull_input(const std::vector<Vector<N, float>>& source_points, std::vector<float>* flat)
{
    if (source_points.empty()) throw ErrorException("no points for convex hull");
    flat->resize(source_points.size() * N);
    for (size_t i = 0; i < source_points.size(); ++i)
        for (size_t j = 0; j < N; ++j) (*flat)[i * N + j] = source_points[i][j];
}

}  // namespace nsp_painter

// compute_convex_hull<N> (convex_hull.h:70-72)
template <size_t N>
void compute_convex_hull(const std::vector<Vector<N, float>>& source_points, std::vector<ConvexHullFacet<N>>* ch_facets,
                         ProgressRatio* /*progress*/)
{
    std::vector<float> flat;
    nsp_painter::hull_input<N>(source_points, &flat);
    nsp_painter::HullHandle hh;
    nsp_painter::check(nsp_convex_hull(nsp_painter::context(), (int)N, flat.data(), (int64_t)source_points.size(),
                                       0, &hh.h));
    int64_t facets = 0;
    nsp_hull_info(hh.h, nullptr, &facets, nullptr, nullptr, nullptr);
    std::vector<int32_t> v((size_t)facets * N);
    std::vector<double> o((size_t)facets * N);
    nsp_hull_copy(hh.h, v.data(), o.data(), nullptr, nullptr);
    ch_facets->clear();
    ch_facets->reserve((size_t)facets);
    for (int64_t f = 0; f < facets; ++f) {
        std::array<int, N> idx;
        vec<N> ortho;
        for (size_t j = 0; j < N; ++j) {
            idx[j] = v[(size_t)f * N + j];
            ortho[j] = o[(size_t)f * N + j];
        }
        ch_facets->emplace_back(idx, ortho);
    }
}

// compute_delaunay<N> (convex_hull.h:67-69)
template <size_t N>
void compute_delaunay(const std::vector<Vector<N, float>>& source_points, std::vector<vec<N>>* points,
                      std::vector<DelaunaySimplex<N>>* simplices, ProgressRatio* /*progress*/)
{
    std::vector<float> flat;
    nsp_painter::hull_input<N>(source_points, &flat);
    nsp_painter::HullHandle hh;
    nsp_painter::check(nsp_delaunay(nsp_painter::context(), (int)N, flat.data(), (int64_t)source_points.size(), 0,
                                    &hh.h));
    int64_t count = 0;
    nsp_hull_info(hh.h, nullptr, &count, nullptr, nullptr, nullptr);
    constexpr size_t K = N + 1;
    std::vector<int32_t> v((size_t)count * K);
    std::vector<double> o((size_t)count * K * N);
    std::vector<double> p(source_points.size() * N);
    nsp_hull_copy(hh.h, v.data(), o.data(), p.data(), nullptr);
    points->assign(source_points.size(), vec<N>{});
    for (size_t i = 0; i < source_points.size(); ++i)
        for (size_t j = 0; j < N; ++j) (*points)[i][j] = p[i * N + j];
    simplices->clear();
    simplices->reserve((size_t)count);
    for (int64_t s = 0; s < count; ++s) {
        std::array<int, K> idx;
        std::array<vec<N>, K> orthos;
        for (size_t r = 0; r < K; ++r) {
            idx[r] = v[(size_t)s * K + r];
            for (size_t j = 0; j < N; ++j) orthos[r][j] = o[((size_t)s * K + r) * N + j];
        }
        simplices->emplace_back(idx, orthos);
    }
}
