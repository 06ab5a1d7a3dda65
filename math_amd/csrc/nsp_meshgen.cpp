// nsp_meshgen.cpp -- deterministic synthetic meshes for the benchmark and test
// configurations (host only).  The reference builds its test meshes as convex
// hulls of random sphere points (src/painter/shapes/test/sphere_mesh.cpp:64-142),
// which needs GMP; these generators make parametric meshes instead:
//   - a torus triangulated on a (u, v) grid with noisy minor radius (n = 3);
//   - the boundary of [-1,1]^n, each facet cut into g^(n-1) cubes and each
//     cube into (n-1)! Kuhn simplices, optionally projected onto S^(n-1),
//     radially perturbed and with a spherical cap removed.
// Random numbers: splitmix64 (Steele, Lea, Flood 2014) and Box-Muller.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "nsp.h"

namespace {

struct SplitMix {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double gaussian()
    {
        double u1 = uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(6.283185307179586 * u2);
    }
};

}  // namespace

extern "C" int nsp_meshgen_torus(double R, double r, int nu, int nv, double noise, uint64_t seed, float* vertices,
                                 int64_t* V, int32_t* facets, int64_t* F)
{
    if (nu < 3 || nv < 3 || !V || !F) return NSP_E_INVALID;
    *V = (int64_t)nu * nv;
    *F = 2 * (int64_t)nu * nv;
    if (!vertices || !facets) return NSP_OK;
    SplitMix rng{seed};
    const double TWO_PI = 6.283185307179586;
    for (int i = 0; i < nu; ++i)
        for (int j = 0; j < nv; ++j) {
            double u = TWO_PI * i / nu, v = TWO_PI * j / nv;
            double rr = r + noise * (2.0 * rng.uniform() - 1.0);
            int64_t k = (int64_t)i * nv + j;
            vertices[k * 3 + 0] = (float)((R + rr * std::cos(v)) * std::cos(u));
            vertices[k * 3 + 1] = (float)((R + rr * std::cos(v)) * std::sin(u));
            vertices[k * 3 + 2] = (float)(rr * std::sin(v));
        }
    int64_t f = 0;
    for (int i = 0; i < nu; ++i)
        for (int j = 0; j < nv; ++j) {
            int i1 = (i + 1) % nu, j1 = (j + 1) % nv;
            int32_t a = i * nv + j, b = i1 * nv + j, c = i1 * nv + j1, d = i * nv + j1;
            facets[f * 3 + 0] = a;
            facets[f * 3 + 1] = b;
            facets[f * 3 + 2] = c;
            ++f;
            facets[f * 3 + 0] = a;
            facets[f * 3 + 1] = c;
            facets[f * 3 + 2] = d;
            ++f;
        }
    return NSP_OK;
}

extern "C" int nsp_meshgen_cube_sphere(int n, int g, int sphere, double sigma, double cap_angle, uint64_t seed,
                                       float* vertices, int64_t* V, int32_t* facets, int64_t* F)
{
    if (n < 3 || n > NSP_MAX_DIM || g < 1 || !V || !F) return NSP_E_INVALID;
    const int m = n - 1;
    int64_t lattice = 1;
    for (int i = 0; i < n; ++i) lattice *= (g + 1);
    if (lattice > ((int64_t)1 << 31)) return NSP_E_INVALID;
    // boundary lattice points -> vertex ids
    std::vector<int32_t> id(lattice, -1);
    int64_t nv = 0;
    std::vector<int> x(n);
    for (int64_t L = 0; L < lattice; ++L) {
        int64_t r = L;
        bool bnd = false;
        for (int i = 0; i < n; ++i) {
            x[i] = (int)(r % (g + 1));
            r /= (g + 1);
            if (x[i] == 0 || x[i] == g) bnd = true;
        }
        if (bnd) id[L] = (int32_t)nv++;
    }
    int64_t cells = 1, perms = 1;
    for (int i = 0; i < m; ++i) {
        cells *= g;
        perms *= (i + 1);
    }
    const int64_t nf_all = 2 * (int64_t)n * cells * perms;
    // vertices
    std::vector<double> pos((size_t)nv * n);
    for (int64_t L = 0; L < lattice; ++L) {
        if (id[L] < 0) continue;
        int64_t r = L;
        for (int i = 0; i < n; ++i) {
            pos[(size_t)id[L] * n + i] = -1.0 + 2.0 * (double)(r % (g + 1)) / g;
            r /= (g + 1);
        }
    }
    SplitMix rng{seed};
    if (sphere) {
        for (int64_t v = 0; v < nv; ++v) {
            double l = 0;
            for (int i = 0; i < n; ++i) l += pos[v * n + i] * pos[v * n + i];
            l = std::sqrt(l);
            double scale = 1.0;
            if (sigma > 0) scale = 1.0 + sigma * rng.gaussian();
            for (int i = 0; i < n; ++i) pos[v * n + i] = pos[v * n + i] / l * scale;
        }
    }
    double capdir[NSP_MAX_DIM] = {0};
    if (cap_angle > 0) {
        double l = 0;
        for (int i = 0; i < n; ++i) {
            capdir[i] = rng.gaussian();
            l += capdir[i] * capdir[i];
        }
        l = std::sqrt(l);
        for (int i = 0; i < n; ++i) capdir[i] /= l;
    }
    const double cap_cos = std::cos(cap_angle);
    // facets
    std::vector<int32_t> out;
    if (facets) out.reserve((size_t)std::min<int64_t>(nf_all, 1 << 26) * n);
    int64_t nf = 0;
    std::vector<int> perm(m), cell(m), p(n);
    std::vector<int> free_axes(m);
    int32_t simplex[NSP_MAX_DIM];
    for (int axis = 0; axis < n; ++axis) {
        int k = 0;
        for (int i = 0; i < n; ++i)
            if (i != axis) free_axes[k++] = i;
        for (int side = 0; side < 2; ++side) {
            for (int64_t c = 0; c < cells; ++c) {
                int64_t r = c;
                for (int i = 0; i < m; ++i) {
                    cell[i] = (int)(r % g);
                    r /= g;
                }
                std::iota(perm.begin(), perm.end(), 0);
                do {
                    // Kuhn simplex: walk from the cell corner along the permuted axes
                    for (int i = 0; i < n; ++i) p[i] = 0;
                    p[axis] = side ? g : 0;
                    for (int i = 0; i < m; ++i) p[free_axes[i]] = cell[i];
                    for (int v = 0; v < n; ++v) {
                        if (v > 0) p[free_axes[perm[v - 1]]] += 1;
                        int64_t L = 0;
                        for (int i = n - 1; i >= 0; --i) L = L * (g + 1) + p[i];
                        simplex[v] = id[L];
                    }
                    bool keep = true;
                    if (cap_angle > 0) {
                        double cen[NSP_MAX_DIM] = {0}, l = 0, d = 0;
                        for (int v = 0; v < n; ++v)
                            for (int i = 0; i < n; ++i) cen[i] += pos[(size_t)simplex[v] * n + i];
                        for (int i = 0; i < n; ++i) l += cen[i] * cen[i];
                        l = std::sqrt(l);
                        for (int i = 0; i < n; ++i) d += cen[i] / l * capdir[i];
                        keep = d < cap_cos;
                    }
                    if (keep) {
                        if (facets) out.insert(out.end(), simplex, simplex + n);
                        ++nf;
                    }
                } while (std::next_permutation(perm.begin(), perm.end()));
            }
        }
    }
    *V = nv;
    *F = nf;
    if (vertices)
        for (int64_t k = 0; k < nv * n; ++k) vertices[k] = (float)pos[k];
    if (facets) std::memcpy(facets, out.data(), sizeof(int32_t) * out.size());
    return NSP_OK;
}
