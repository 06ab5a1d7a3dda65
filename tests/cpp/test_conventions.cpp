// test_conventions.cpp -- prints what the C++17 drop-in painter.h derives from
// its inputs, for the Python tests to compare against the reference's rules:
//
//   test_conventions paintbrush <paint_height> <size0> <size1> ...
//       VisibleBarPaintbrush (src/painter/paintbrushes/paintbrush.h:26-233)
//       over a 2-, 3- or 4-dimensional screen: "passes P pixels Q" after one
//       pass of next_pixel calls, then the pixels in the order handed out.
//       Host only (no GPU call).
//   test_conventions scene <g> <min_screen> <max_screen>
//       single_object_scene(background, colour, diffuse, min, max, mesh)
//       (src/painter/scenes/single_object.h:102-179) for the 4-space mesh of
//       the library's cube-sphere generator with g cells per edge: the mesh
//       min/max, the parallel projector's camera, direction, axes, units per
//       pixel and screen, and the light (kind, position, colour).  Builds the
//       mesh on the GPU.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "painter/painter.h"

namespace {

template <size_t N>
int paintbrush(int h, const std::vector<int>& sizes)
{
    std::array<int, N> sz{};
    for (size_t i = 0; i < N; ++i) sz[i] = sizes[i];
    VisibleBarPaintbrush<N> brush(sz, h, 1);
    brush.first_pass();
    std::vector<std::array<int_least16_t, N>> order;
    std::array<int_least16_t, N> p{};
    while (brush.next_pixel(0, 0, &p)) order.push_back(p);
    long long passes, pixels, rays, samples;
    double dur;
    brush.statistics(&passes, &pixels, &rays, &samples, &dur);
    std::printf("passes %lld pixels %lld\n", passes, pixels);
    for (const auto& q : order) {
        for (size_t i = 0; i < N; ++i) std::printf(i ? " %d" : "%d", (int)q[i]);
        std::printf("\n");
    }
    return 0;
}

int scene(int g, int min_screen, int max_screen)
{
    int64_t V = 0, F = 0;
    nsp_meshgen_cube_sphere(4, g, 1, 0.03, 0.0, 7, nullptr, &V, nullptr, &F);
    std::vector<float> v(V * 4);
    std::vector<int32_t> f(F * 4);
    nsp_meshgen_cube_sphere(4, g, 1, 0.03, 0.0, 7, v.data(), &V, f.data(), &F);
    Obj<4> obj;
    for (int64_t i = 0; i < V; ++i) obj.vertices.push_back({v[i * 4], v[i * 4 + 1], v[i * 4 + 2], v[i * 4 + 3]});
    for (int64_t i = 0; i < F; ++i) {
        Obj<4>::Facet fc;
        for (int k = 0; k < 4; ++k) fc.vertices[k] = f[i * 4 + k];
        obj.facets.push_back(fc);
    }
    auto mesh = std::make_shared<const Mesh<4, double>>(&obj, identity_matrix<5, double>(), 1, nullptr);
    auto sc = single_object_scene<4, double>(Color(0.1f, 0.2f, 0.3f), Color(1.0f), 0.8f, min_screen, max_screen, mesh);
    Vector<4, double> mn, mx;
    mesh->min_max(&mn, &mx);
    const nsp_projector pr = sc->projector().data();
    std::printf("min %.17g %.17g %.17g %.17g\n", mn[0], mn[1], mn[2], mn[3]);
    std::printf("max %.17g %.17g %.17g %.17g\n", mx[0], mx[1], mx[2], mx[3]);
    std::printf("projector %d\n", pr.kind);
    std::printf("camera_org %.17g %.17g %.17g %.17g\n", pr.camera_org[0], pr.camera_org[1], pr.camera_org[2],
                pr.camera_org[3]);
    std::printf("camera_dir %.17g %.17g %.17g %.17g\n", pr.camera_dir[0], pr.camera_dir[1], pr.camera_dir[2],
                pr.camera_dir[3]);
    for (int a = 0; a < 3; ++a)
        std::printf("axis %.17g %.17g %.17g %.17g\n", pr.screen_axes[a][0], pr.screen_axes[a][1],
                    pr.screen_axes[a][2], pr.screen_axes[a][3]);
    std::printf("units_per_pixel %.17g\n", pr.units_per_pixel);
    std::printf("screen %d %d %d\n", pr.screen_size[0], pr.screen_size[1], pr.screen_size[2]);
    for (const LightSource<4, double>* l : sc->light_sources()) {
        const nsp_light d = l->data();
        std::printf("light %d %.17g %.17g %.17g %.17g %.9g %.9g %.9g\n", d.kind, d.position[0], d.position[1],
                    d.position[2], d.position[3], d.rgb[0], d.rgb[1], d.rgb[2]);
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    try {
        if (mode == "paintbrush" && argc >= 5) {
            const int h = std::atoi(argv[2]);
            std::vector<int> sizes;
            for (int i = 3; i < argc; ++i) sizes.push_back(std::atoi(argv[i]));
            switch (sizes.size()) {
            case 2: return paintbrush<2>(h, sizes);
            case 3: return paintbrush<3>(h, sizes);
            case 4: return paintbrush<4>(h, sizes);
            default: return 2;
            }
        }
        if (mode == "scene" && argc == 5) return scene(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]));
    } catch (const std::exception& e) {
        std::printf("FAIL %s\n", e.what());
        return 1;
    }
    return 2;
}
