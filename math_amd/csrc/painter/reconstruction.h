// reconstruction.h -- C++17 drop-in for Math Viewer's Cocone and BoundCocone
// manifold constructors (src/geometry/cocone/reconstruction.h:28-57) on
// libnsp: the same class and function names and argument meaning, errors
// thrown as ErrorException with the reference's messages.  The Delaunay
// simplices use exact integer predicates; the Voronoi vertices, poles, cells
// and cocone flags are computed on the GPU of the painter's context.
#pragma once

#include <array>
#include <cstdint>
#include <memory>
#include <vector>

#include "convex_hull.h"

namespace nsp_painter {

struct ManifoldHandle {
    nsp_manifold* m = nullptr;
    ManifoldHandle() = default;
    ManifoldHandle(const ManifoldHandle&) = delete;
    ManifoldHandle& operator=(const ManifoldHandle&) = delete;
    ~ManifoldHandle() { nsp_manifold_destroy(m); }
};

template <size_t N>
class ManifoldConstructorImpl {
    ManifoldHandle m_h;
    int64_t m_points = 0;

    void result(int64_t k, std::vector<vec<N>>* normals, std::vector<std::array<int, N>>* facets) const
    {
        std::vector<double> nd((size_t)m_points * N);
        std::vector<int32_t> fd((size_t)k * N);
        check(nsp_manifold_result(m_h.m, nd.data(), fd.data()));
        normals->assign((size_t)m_points, vec<N>{});
        for (int64_t i = 0; i < m_points; ++i)
            for (size_t j = 0; j < N; ++j) (*normals)[i][j] = nd[(size_t)i * N + j];
        facets->resize((size_t)k);
        for (int64_t f = 0; f < k; ++f)
            for (size_t j = 0; j < N; ++j) (*facets)[f][j] = fd[(size_t)f * N + j];
    }

public:
    ManifoldConstructorImpl(const std::vector<Vector<N, float>>& source_points, bool cocone_only)
    {
        std::vector<float> flat(source_points.size() * N);
        for (size_t i = 0; i < source_points.size(); ++i)
            for (size_t j = 0; j < N; ++j) flat[i * N + j] = source_points[i][j];
        check(nsp_manifold_create(context(), (int)N, flat.data(), (int64_t)source_points.size(), cocone_only ? 1 : 0,
                                  &m_h.m));
        nsp_manifold_info(m_h.m, nullptr, &m_points, nullptr, nullptr);
    }

    std::vector<std::array<int, N + 1>> delaunay_objects() const
    {
        int64_t S = 0;
        nsp_manifold_info(m_h.m, nullptr, nullptr, &S, nullptr);
        std::vector<int32_t> v((size_t)S * (N + 1));
        check(nsp_manifold_objects(m_h.m, v.data()));
        std::vector<std::array<int, N + 1>> out((size_t)S);
        for (int64_t s = 0; s < S; ++s)
            for (size_t j = 0; j <= N; ++j) out[s][j] = v[(size_t)s * (N + 1) + j];
        return out;
    }

    void cocone(std::vector<vec<N>>* normals, std::vector<std::array<int, N>>* facets) const
    {
        int64_t k = 0;
        check(nsp_manifold_cocone(context(), m_h.m, &k));
        result(k, normals, facets);
    }

    void bound_cocone(double rho, double alpha, std::vector<vec<N>>* normals,
                      std::vector<std::array<int, N>>* facets) const
    {
        int64_t k = 0;
        check(nsp_manifold_bound_cocone(context(), m_h.m, rho, alpha, &k));
        result(k, normals, facets);
    }
};

}  // namespace nsp_painter

// ManifoldConstructorCocone<N> (reconstruction.h:28-37)
template <size_t N>
struct ManifoldConstructorCocone {
    virtual ~ManifoldConstructorCocone() = default;
    virtual std::vector<std::array<int, N + 1>> delaunay_objects() const = 0;
    virtual void cocone(std::vector<vec<N>>* vertex_normals, std::vector<std::array<int, N>>* cocone_triangles,
                        ProgressRatio* progress) const = 0;
};

// ManifoldConstructor<N> (reconstruction.h:39-48)
template <size_t N>
struct ManifoldConstructor {
    virtual ~ManifoldConstructor() = default;
    virtual std::vector<std::array<int, N + 1>> delaunay_objects() const = 0;
    virtual void cocone(std::vector<vec<N>>* normals, std::vector<std::array<int, N>>* facets,
                        ProgressRatio* progress) const = 0;
    virtual void bound_cocone(double rho, double alpha, std::vector<vec<N>>* normals,
                              std::vector<std::array<int, N>>* facets, ProgressRatio* progress) const = 0;
};

namespace nsp_painter {

template <size_t N>
class ManifoldConstructorGpu final : public ManifoldConstructor<N>, public ManifoldConstructorCocone<N> {
    ManifoldConstructorImpl<N> m_impl;

public:
    ManifoldConstructorGpu(const std::vector<Vector<N, float>>& points, bool cocone_only) : m_impl(points, cocone_only)
    {
    }
    std::vector<std::array<int, N + 1>> delaunay_objects() const override { return m_impl.delaunay_objects(); }
    void cocone(std::vector<vec<N>>* normals, std::vector<std::array<int, N>>* facets, ProgressRatio*) const override
    {
        m_impl.cocone(normals, facets);
    }
    void bound_cocone(double rho, double alpha, std::vector<vec<N>>* normals, std::vector<std::array<int, N>>* facets,
                      ProgressRatio*) const override
    {
        m_impl.bound_cocone(rho, alpha, normals, facets);
    }
};

}  // namespace nsp_painter

// create_manifold_constructor<N> (reconstruction.h:50-52)
template <size_t N>
std::unique_ptr<ManifoldConstructor<N>> create_manifold_constructor(const std::vector<Vector<N, float>>& source_points,
                                                                    ProgressRatio* /*progress*/)
{
    return std::make_unique<nsp_painter::ManifoldConstructorGpu<N>>(source_points, false);
}

// create_manifold_constructor_cocone<N> (reconstruction.h:54-56)
template <size_t N>
std::unique_ptr<ManifoldConstructorCocone<N>> create_manifold_constructor_cocone(
    const std::vector<Vector<N, float>>& source_points, ProgressRatio* /*progress*/)
{
    return std::make_unique<nsp_painter::ManifoldConstructorGpu<N>>(source_points, true);
}
