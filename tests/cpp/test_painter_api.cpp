// test_painter_api.cpp -- the reference's headless painter test
// (src/painter/test/test_painter.cpp:138-233: N=4, double, sphere mesh,
// single_object_scene, VisibleBarPaintbrush with paint height 2 and one pass)
// run through the C++17 drop-in painter.h on the GPU.
//
// Usage: test_painter_api <n_cells_per_edge> <max_screen> <spp>
// Prints "OK <pixels> <rays> <mean luminance>" or "FAIL <reason>".
#include <cstdio>
#include <cstdlib>
#include <map>
#include <set>

#include "painter/painter.h"

namespace {

constexpr Srgb8 BACKGROUND_COLOR(50, 100, 150);
constexpr Srgb8 DEFAULT_COLOR(150, 170, 150);

// Obj<4> of the tesseract boundary cut into g^3 cubes per facet and 6 Kuhn
// tetrahedra per cube, projected onto S^3 (the library's generator).
Obj<4> sphere_obj(int g)
{
    int64_t V = 0, F = 0;
    nsp_meshgen_cube_sphere(4, g, 1, 0.0, 0.0, 1, nullptr, &V, nullptr, &F);
    std::vector<float> v(V * 4);
    std::vector<int32_t> f(F * 4);
    nsp_meshgen_cube_sphere(4, g, 1, 0.0, 0.0, 1, v.data(), &V, f.data(), &F);
    Obj<4> obj;
    for (int64_t i = 0; i < V; ++i) obj.vertices.push_back({v[i * 4], v[i * 4 + 1], v[i * 4 + 2], v[i * 4 + 3]});
    for (int64_t i = 0; i < F; ++i) {
        Obj<4>::Facet fc;
        for (int k = 0; k < 4; ++k) fc.vertices[k] = f[i * 4 + k];
        obj.facets.push_back(fc);
    }
    return obj;
}

class Images final : public PainterNotifier<3> {
public:
    std::map<std::array<int_least16_t, 3>, Color> pixels;
    int before = 0, after = 0, errors = 0;
    std::string last_error;
    void painter_pixel_before(const std::array<int_least16_t, 3>&) noexcept override { ++before; }
    void painter_pixel_after(const std::array<int_least16_t, 3>& p, const Color& c) noexcept override
    {
        ++after;
        pixels[p] = c;
    }
    void painter_error_message(const std::string& msg) noexcept override
    {
        ++errors;
        last_error = msg;
    }
};

}  // namespace

int main(int argc, char** argv)
{
    int g = argc > 1 ? std::atoi(argv[1]) : 4;
    int max_screen = argc > 2 ? std::atoi(argv[2]) : 16;
    int spp = argc > 3 ? std::atoi(argv[3]) : 4;
    try {
        Obj<4> obj = sphere_obj(g);
        auto mesh = std::make_shared<const Mesh<4, double>>(&obj, identity_matrix<5, double>(), 1, nullptr);
        std::unique_ptr<const PaintObjects<4, double>> scene =
            single_object_scene<4, double>(BACKGROUND_COLOR, DEFAULT_COLOR, 1, 10, max_screen, mesh);
        const auto& size = scene->projector().screen_size();
        VisibleBarPaintbrush<3> brush(size, 2, 1);
        Images images;
        std::atomic_bool stop = false;
        paint(&images, spp, *scene, &brush, 8, &stop, false);
        if (images.errors) {
            std::printf("FAIL painter error: %s\n", images.last_error.c_str());
            return 1;
        }
        size_t expect = (size_t)size[0] * size[1] * size[2];
        if (images.pixels.size() != expect || (size_t)images.after != expect || images.before != images.after) {
            std::printf("FAIL pixel count %zu (after %d, before %d), expected %zu\n", images.pixels.size(),
                        images.after, images.before, expect);
            return 1;
        }
        double lum = 0;
        for (auto& [p, c] : images.pixels) {
            if (!std::isfinite(c.r) || !std::isfinite(c.g) || !std::isfinite(c.b) || c.r < 0) {
                std::printf("FAIL bad colour\n");
                return 1;
            }
            lum += c.luminance();
        }
        long long passes, pixels, rays, samples;
        double dur;
        brush.statistics(&passes, &pixels, &rays, &samples, &dur);
        // error path: thread count < 1 is reported, not thrown (painter.cpp:521-527)
        Images bad;
        VisibleBarPaintbrush<3> brush2(size, 2, 1);
        paint(&bad, spp, *scene, &brush2, 0, &stop, false);
        if (bad.errors != 1 || bad.last_error.find("thread count") == std::string::npos) {
            std::printf("FAIL thread-count error not reported\n");
            return 1;
        }
        // the other projectors through the explicit-projector single_object_scene
        // (single_object.h:87-100): camera 3.5 units up the last axis, looking down it
        for (int kind = 0; kind < 2; ++kind) {
            Vector<4, double> org{0.0, 0.0, 0.0, 3.5}, dir{0.0, 0.0, 0.0, -1.0};
            std::array<Vector<4, double>, 3> axes = {Vector<4, double>{1.0, 0.0, 0.0, 0.0},
                                                     Vector<4, double>{0.0, 1.0, 0.0, 0.0},
                                                     Vector<4, double>{0.0, 0.0, 1.0, 0.0}};
            std::array<int, 3> sz = {max_screen, max_screen, max_screen};
            std::unique_ptr<const Projector<4, double>> proj;
            if (kind == 0)
                proj = std::make_unique<const VisiblePerspectiveProjector<4, double>>(org, dir, axes, 50.0, sz);
            else
                proj = std::make_unique<const VisibleSphericalProjector<4, double>>(org, dir, axes, 50.0, sz);
            auto light = std::make_unique<const VisiblePointLight<4, double>>(Vector<4, double>{2.0, 2.0, 2.0, 2.0},
                                                                              Color(1), 4.0);
            auto sc = single_object_scene<4, double>(BACKGROUND_COLOR, DEFAULT_COLOR, 1, std::move(proj),
                                                     std::move(light), mesh);
            Images im;
            VisibleBarPaintbrush<3> b3(sz, 2, 1);
            paint(&im, spp, *sc, &b3, 8, &stop, false);
            size_t want = (size_t)sz[0] * sz[1] * sz[2];
            const Color bg = BACKGROUND_COLOR;
            size_t object = 0;
            for (auto& [p, c] : im.pixels) object += std::fabs(c.r - bg.r) + std::fabs(c.b - bg.b) > 1e-3f;
            if (im.errors || im.pixels.size() != want || object == 0 || object == want) {
                std::printf("FAIL projector %d: errors %d (%s), pixels %zu, object pixels %zu\n", kind, im.errors,
                            im.last_error.c_str(), im.pixels.size(), object);
                return 1;
            }
        }
        bool threw = false;
        try {
            VisiblePerspectiveProjector<4, double> p(Vector<4, double>{0.0, 0.0, 0.0, 3.5},
                                                     Vector<4, double>{0.0, 0.0, 0.0, -1.0},
                                                     {Vector<4, double>{1.0, 0.0, 0.0, 0.0},
                                                      Vector<4, double>{0.0, 1.0, 0.0, 0.0},
                                                      Vector<4, double>{0.0, 0.0, 1.0, 0.0}},
                                                     180.0, {4, 4, 4});
        } catch (const std::exception&) {
            threw = true;
        }
        if (!threw) {
            std::printf("FAIL perspective view angle 180 accepted\n");
            return 1;
        }
        std::printf("OK %zu %lld %.6f %lld\n", images.pixels.size(), rays, lum / expect, passes);
        return 0;
    } catch (const std::exception& e) {
        std::printf("FAIL exception %s\n", e.what());
        return 1;
    }
}
