// painter.h -- C++17 drop-in for the painter API of Math Viewer
// (templeblock/math src/painter/), running on libnsp (include/nsp.h).
//
// The reference API is a set of virtual interfaces evaluated per ray on CPU
// threads (src/painter/objects.h:143-223, src/painter/painter.h:26-42).  Here
// the concrete classes callers construct -- Mesh, VisibleSharedMesh,
// VisiblePointLight, VisibleConstantLight, VisibleParallelProjector,
// VisiblePerspectiveProjector, VisibleSphericalProjector,
// VisibleBarPaintbrush, single_object_scene -- carry their data, and
// paint<N,T>() hands that data to the GPU through the C ABI.  Names,
// argument meaning and error behaviour follow the reference:
//   - paint() never throws; failures go to PainterNotifier::painter_error_message
//     with the "Painter error:\n" prefix (src/painter/painter.cpp:599-625);
//   - pixels are delivered through painter_pixel_before/after in the
//     paintbrush's order, one pass at a time (painter.cpp:410-448), with the
//     running mean colour of each pixel;
//   - `stop` is polled between wavefront batches and passes.
// T selects the host-side precision of the scene description; the trace
// runs in FP32 with FP64 near-boundary re-tests (set_painter_fp64(true) for
// FP64 everywhere).
#pragma once

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <limits>
#include <cstdint>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "nsp.h"

template <size_t N, typename T>
using Vector = std::array<T, N>;

// RGB colour with float components (src/com/color/color.h:35-111).
struct Color {
    float r = 0, g = 0, b = 0;
    Color() = default;
    explicit Color(float v) : r(v), g(v), b(v) {}
    Color(float r_, float g_, float b_) : r(r_), g(g_), b(b_) {}
    float luminance() const { return 0.2126f * r + 0.7152f * g + 0.0722f * b; }
};

// sRGB 8-bit colour (src/com/color/color.h:26-33), converted with srgb_to_rgb
// (src/com/color/conversion.cpp:188-205).
struct Srgb8 {
    unsigned char red, green, blue;
    constexpr Srgb8(unsigned char r, unsigned char g, unsigned char b) : red(r), green(g), blue(b) {}
    static float to_linear(unsigned char c8)
    {
        float c = c8 / 255.0f;
        if (c >= 1) return 1;
        if (c >= 0.04045f) return std::pow((c + 0.055f) / 1.055f, 2.4f);
        if (c > 0) return c / 12.92f;
        return 0;
    }
    operator Color() const { return Color(to_linear(red), to_linear(green), to_linear(blue)); }
};

class ErrorException : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

namespace nsp_painter {

// One library context per process on device 0.
inline nsp_context* context()
{
    static std::once_flag once;
    static nsp_context* ctx = nullptr;
    static std::string err;
    std::call_once(once, [] {
        if (nsp_context_create(0, &ctx) != NSP_OK) err = nsp_last_error(nullptr);
    });
    if (!ctx) throw ErrorException("GPU context: " + err);
    return ctx;
}

inline void check(int rc)
{
    if (rc != NSP_OK) throw ErrorException(nsp_last_error(context()));
}

inline bool& fp64_flag()
{
    static bool v = false;
    return v;
}

}  // namespace nsp_painter

inline void set_painter_fp64(bool on) { nsp_painter::fp64_flag() = on; }

// Obj<N> (src/obj/obj.h:27-82): the mesh input.
template <size_t N>
struct Obj {
    struct Facet {
        std::array<int, N> vertices;
        std::array<int, N> normals;
        bool has_normal = false;
    };
    std::vector<Vector<N, float>> vertices;
    std::vector<Vector<N, float>> normals;
    std::vector<Facet> facets;
};

class ProgressRatio {
public:
    explicit ProgressRatio(void*) {}
};

template <size_t N, typename T>
using Matrix = std::array<std::array<T, N>, N>;

template <size_t N, typename T>
Matrix<N, T> identity_matrix()
{
    Matrix<N, T> m{};
    for (size_t i = 0; i < N; ++i) m[i][i] = 1;
    return m;
}

// Mesh<N,T> (src/painter/shapes/mesh.h:32-87): the FP64 basis and the
// 2^N-tree are built on the GPU by the constructor.
template <size_t N, typename T>
class Mesh {
    nsp_mesh* m_mesh = nullptr;
    Vector<N, T> m_min, m_max;

public:
    Mesh(const Obj<N>* obj, const Matrix<N + 1, T>& vertex_matrix, unsigned /*thread_count*/, ProgressRatio*)
    {
        if (obj->vertices.empty()) throw ErrorException("No vertices found in obj");
        if (obj->facets.empty()) throw ErrorException("No facets found in obj");
        std::vector<float> v(obj->vertices.size() * N), nrm(obj->normals.size() * N);
        for (size_t i = 0; i < obj->vertices.size(); ++i)
            for (size_t k = 0; k < N; ++k) v[i * N + k] = obj->vertices[i][k];
        for (size_t i = 0; i < obj->normals.size(); ++i)
            for (size_t k = 0; k < N; ++k) nrm[i * N + k] = obj->normals[i][k];
        std::vector<int32_t> f(obj->facets.size() * N), fn(obj->facets.size() * N);
        bool any_normals = false;
        for (size_t i = 0; i < obj->facets.size(); ++i)
            for (size_t k = 0; k < N; ++k) {
                f[i * N + k] = obj->facets[i].vertices[k];
                fn[i * N + k] = obj->facets[i].has_normal ? obj->facets[i].normals[k] : -1;
                any_normals = any_normals || obj->facets[i].has_normal;
            }
        double mtx[(N + 1) * (N + 1)];
        for (size_t r = 0; r <= N; ++r)
            for (size_t c = 0; c <= N; ++c) mtx[r * (N + 1) + c] = (double)vertex_matrix[r][c];
        bool use_n = any_normals && !obj->normals.empty();
        nsp_painter::check(nsp_mesh_create(nsp_painter::context(), (int)N, v.data(), (int64_t)obj->vertices.size(),
                                           f.data(), (int64_t)obj->facets.size(), use_n ? nrm.data() : nullptr,
                                           use_n ? (int64_t)obj->normals.size() : 0, use_n ? fn.data() : nullptr, mtx,
                                           &m_mesh));
        double mn[NSP_MAX_DIM], mx[NSP_MAX_DIM];
        nsp_mesh_min_max(m_mesh, mn, mx);
        for (size_t i = 0; i < N; ++i) {
            m_min[i] = (T)mn[i];
            m_max[i] = (T)mx[i];
        }
    }
    Mesh(const Mesh&) = delete;
    Mesh& operator=(const Mesh&) = delete;
    ~Mesh() { nsp_mesh_destroy(m_mesh); }
    void min_max(Vector<N, T>* min, Vector<N, T>* max) const
    {
        *min = m_min;
        *max = m_max;
    }
    nsp_mesh* handle() const { return m_mesh; }
};

// SurfaceProperties (src/painter/objects.h:27-139), the parts the painter reads.
template <size_t N, typename T>
class SurfaceProperties {
    Color m_color{0}, m_light_source_color{0};
    T m_diffuse = 1, m_fresnel = 0;
    bool m_light_source = false;

public:
    void set_color(const Color& c) { m_color = c; }
    const Color& get_color() const { return m_color; }
    void set_light_source_color(const Color& c) { m_light_source_color = c; }
    const Color& get_light_source_color() const { return m_light_source_color; }
    void set_diffuse_and_fresnel(T d, T f)
    {
        m_diffuse = d;
        m_fresnel = f;
    }
    T get_diffuse() const { return m_diffuse; }
    void set_light_source(bool v) { m_light_source = v; }
    bool is_light_source() const { return m_light_source; }
    nsp_surface to_nsp() const
    {
        nsp_surface s{};
        s.rgb[0] = m_color.r;
        s.rgb[1] = m_color.g;
        s.rgb[2] = m_color.b;
        s.diffuse = (float)m_diffuse;
        s.light_source = m_light_source ? 1 : 0;
        s.light_rgb[0] = m_light_source_color.r;
        s.light_rgb[1] = m_light_source_color.g;
        s.light_rgb[2] = m_light_source_color.b;
        return s;
    }
};

template <size_t N, typename T>
struct GenericObject {
    virtual ~GenericObject() = default;
    virtual void min_max(Vector<N, T>* min, Vector<N, T>* max) const = 0;
};

template <size_t N, typename T>
struct LightSource {
    virtual ~LightSource() = default;
    virtual nsp_light data() const = 0;
};

template <size_t N, typename T>
struct Projector {
    virtual ~Projector() = default;
    virtual const std::array<int, N - 1>& screen_size() const = 0;
    virtual nsp_projector data() const = 0;
};

// VisibleSharedMesh (src/painter/visible_shapes.h:119-168)
template <size_t N, typename T>
class VisibleSharedMesh final : public GenericObject<N, T>, public SurfaceProperties<N, T> {
    std::shared_ptr<const Mesh<N, T>> m_mesh;

public:
    VisibleSharedMesh(const std::shared_ptr<const Mesh<N, T>>& mesh) : m_mesh(mesh) {}
    void min_max(Vector<N, T>* min, Vector<N, T>* max) const override { m_mesh->min_max(min, max); }
    const Mesh<N, T>& mesh() const { return *m_mesh; }
};

// VisiblePointLight (src/painter/visible_lights.h:24-39, PointLight light_source.h:26-57)
template <size_t N, typename T>
class VisiblePointLight final : public LightSource<N, T> {
    nsp_light m_light{};

public:
    VisiblePointLight(const Vector<N, T>& location, const Color& color, T unit_intensity_distance)
    {
        m_light.kind = NSP_LIGHT_POINT;
        for (size_t i = 0; i < N; ++i) m_light.position[i] = (double)location[i];
        m_light.rgb[0] = color.r;
        m_light.rgb[1] = color.g;
        m_light.rgb[2] = color.b;
        m_light.unit_intensity_distance = (double)unit_intensity_distance;
    }
    nsp_light data() const override { return m_light; }
};

// VisibleConstantLight (src/painter/visible_lights.h:41-55)
template <size_t N, typename T>
class VisibleConstantLight final : public LightSource<N, T> {
    nsp_light m_light{};

public:
    VisibleConstantLight(const Vector<N, T>& location, const Color& color)
    {
        m_light.kind = NSP_LIGHT_CONSTANT;
        for (size_t i = 0; i < N; ++i) m_light.position[i] = (double)location[i];
        m_light.rgb[0] = color.r;
        m_light.rgb[1] = color.g;
        m_light.rgb[2] = color.b;
        m_light.unit_intensity_distance = 1;
    }
    nsp_light data() const override { return m_light; }
};

namespace nsp_painter {
// the camera, axes and screen of a projector (projector.h:51-86)
template <size_t N, typename T>
nsp_projector projector_data(int kind, const Vector<N, T>& camera_org, const Vector<N, T>& camera_dir,
                             const std::array<Vector<N, T>, N - 1>& screen_axes,
                             const std::array<int, N - 1>& screen_size)
{
    nsp_projector p{};
    p.kind = kind;
    for (size_t i = 0; i < N; ++i) {
        p.camera_org[i] = (double)camera_org[i];
        p.camera_dir[i] = (double)camera_dir[i];
    }
    for (size_t a = 0; a + 1 < N; ++a) {
        for (size_t i = 0; i < N; ++i) p.screen_axes[a][i] = (double)screen_axes[a][i];
        p.screen_size[a] = screen_size[a];
        if (screen_size[a] < 1)
            throw ErrorException("Projection size " + std::to_string(a) + " is not positive (" +
                                 std::to_string(screen_size[a]) + ")");
    }
    return p;
}
}  // namespace nsp_painter

// VisiblePerspectiveProjector (src/painter/visible_projectors.h:24-46, PerspectiveProjector
// projector.h:102-151)
template <size_t N, typename T>
class VisiblePerspectiveProjector final : public Projector<N, T> {
    nsp_projector m_p;
    std::array<int, N - 1> m_size;

public:
    VisiblePerspectiveProjector(const Vector<N, T>& camera_org, const Vector<N, T>& camera_dir,
                                const std::array<Vector<N, T>, N - 1>& screen_axes, T width_view_angle_degrees,
                                const std::array<int, N - 1>& screen_size)
        : m_p(nsp_painter::projector_data<N, T>(NSP_PROJECTOR_PERSPECTIVE, camera_org, camera_dir, screen_axes,
                                                screen_size)),
          m_size(screen_size)
    {
        if (!(width_view_angle_degrees > 0 && width_view_angle_degrees < 180))
            throw ErrorException("Perspective projection: error view angle " +
                                 std::to_string((double)width_view_angle_degrees));
        m_p.view_angle_degrees = (double)width_view_angle_degrees;
    }
    const std::array<int, N - 1>& screen_size() const override { return m_size; }
    nsp_projector data() const override { return m_p; }
};

// VisibleParallelProjector (src/painter/visible_projectors.h:48-70)
template <size_t N, typename T>
class VisibleParallelProjector final : public Projector<N, T> {
    nsp_projector m_p;
    std::array<int, N - 1> m_size;

public:
    VisibleParallelProjector(const Vector<N, T>& camera_org, const Vector<N, T>& camera_dir,
                             const std::array<Vector<N, T>, N - 1>& screen_axes, T units_per_pixel,
                             const std::array<int, N - 1>& screen_size)
        : m_p(nsp_painter::projector_data<N, T>(NSP_PROJECTOR_PARALLEL, camera_org, camera_dir, screen_axes,
                                                screen_size)),
          m_size(screen_size)
    {
        if (!(units_per_pixel > 0)) throw ErrorException("Error units per pixel for parallel projection");
        m_p.units_per_pixel = (double)units_per_pixel;
    }
    const std::array<int, N - 1>& screen_size() const override { return m_size; }
    nsp_projector data() const override { return m_p; }
};

// VisibleSphericalProjector (src/painter/visible_projectors.h:72-94, SphericalProjector
// projector.h:204-275)
template <size_t N, typename T>
class VisibleSphericalProjector final : public Projector<N, T> {
    nsp_projector m_p;
    std::array<int, N - 1> m_size;

public:
    VisibleSphericalProjector(const Vector<N, T>& camera_org, const Vector<N, T>& camera_dir,
                              const std::array<Vector<N, T>, N - 1>& screen_axes, T width_view_angle_degrees,
                              const std::array<int, N - 1>& screen_size)
        : m_p(nsp_painter::projector_data<N, T>(NSP_PROJECTOR_SPHERICAL, camera_org, camera_dir, screen_axes,
                                                screen_size)),
          m_size(screen_size)
    {
        const double angle = (double)width_view_angle_degrees;
        const double sin_alpha = std::sin(angle * 0.5 * 3.14159265358979323846 / 180);
        if (!(angle > 0))
            throw ErrorException("Spherical projection view angle " + std::to_string(angle) + " is not positive");
        const double k = sin_alpha / screen_size[0];
        double r = sin_alpha * sin_alpha;
        for (size_t i = 1; i + 1 < N; ++i) r += (k * screen_size[i]) * (k * screen_size[i]);
        if (!(r < 1))
            throw ErrorException("Spherical projection view angle " + std::to_string(angle) + " is too big");
        m_p.view_angle_degrees = angle;
    }
    const std::array<int, N - 1>& screen_size() const override { return m_size; }
    nsp_projector data() const override { return m_p; }
};

// Paintbrush interface (src/painter/objects.h:190-209)
template <size_t N>
struct Paintbrush {
    virtual ~Paintbrush() = default;
    virtual void first_pass() noexcept = 0;
    virtual bool next_pixel(int previous_pixel_ray_count, int previous_pixel_sample_count,
                            std::array<int_least16_t, N>* pixel) noexcept = 0;
    virtual bool next_pass() noexcept = 0;
    virtual const std::array<int, N>& screen_size() const noexcept = 0;
    virtual void statistics(long long* pass_count, long long* pixel_count, long long* ray_count,
                            long long* sample_count, double* previous_pass_duration) const noexcept = 0;
};

// VisibleBarPaintbrush / BarPaintbrush (src/painter/paintbrushes/paintbrush.h:26-233):
// bars of paint_height over all dimensions but the last, the last dimension
// innermost-but-one, the second coordinate flipped for 2D screens.
template <size_t N>
class VisibleBarPaintbrush final : public Paintbrush<N> {
    using Pixel = std::array<int_least16_t, N>;
    std::array<int, N> m_screen_size;
    std::vector<Pixel> m_pixels;
    int m_max_pass_count;
    size_t m_current = 0;
    long long m_pass_count = 1, m_pixel_count = 0, m_ray_count = 0, m_sample_count = 0;
    double m_prev = 0, m_start = -1;
    mutable std::mutex m_lock;

    static double now()
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    }

    void generate(const std::array<int, N>& sizes, int h)
    {
        // sizes are in reversed order here (last screen dimension first)
        std::array<int, N> lo{}, hi{};
        Pixel p{};
        gen_outer(0, sizes, h, lo, hi, p);
    }
    void gen_outer(size_t level, const std::array<int, N>& sizes, int h, std::array<int, N>& lo,
                   std::array<int, N>& hi, Pixel& p)
    {
        if (level < N - 1) {
            for (int i = 0; i < sizes[level]; i += h) {
                lo[level] = i;
                hi[level] = std::min(sizes[level], i + h);
                gen_outer(level + 1, sizes, h, lo, hi, p);
            }
            return;
        }
        for (int i = 0; i < sizes[N - 1]; ++i) {
            p[N - 1] = (int_least16_t)i;
            gen_inner(0, lo, hi, p);
        }
    }
    void gen_inner(size_t level, const std::array<int, N>& lo, const std::array<int, N>& hi, Pixel& p)
    {
        if (level == N - 1) {
            m_pixels.push_back(p);
            return;
        }
        for (int i = lo[level]; i < hi[level]; ++i) {
            p[level] = (int_least16_t)i;
            gen_inner(level + 1, lo, hi, p);
        }
    }

public:
    VisibleBarPaintbrush(const std::array<int, N>& screen_size, int paint_height, int max_pass_count)
    {
        for (size_t i = 0; i < N; ++i)
            if (screen_size[i] < 1)
                throw ErrorException("Paintbrush size " + std::to_string(i) + " is not positive (" +
                                     std::to_string(screen_size[i]) + ")");
        if (paint_height < 1) throw ErrorException("Error paintbrush paint height " + std::to_string(paint_height));
        if (!(max_pass_count == -1 || max_pass_count > 0))
            throw ErrorException("Error paintbrush max pass count " + std::to_string(max_pass_count));
        m_screen_size = screen_size;
        m_max_pass_count = max_pass_count;
        std::array<int, N> rev = screen_size;
        std::reverse(rev.begin(), rev.end());
        generate(rev, paint_height);
        for (Pixel& px : m_pixels) {
            std::reverse(px.begin(), px.end());
            if (N == 2) px[1] = (int_least16_t)(m_screen_size[1] - 1 - px[1]);
        }
    }
    const std::array<int, N>& screen_size() const noexcept override { return m_screen_size; }
    void first_pass() noexcept override
    {
        std::lock_guard<std::mutex> lg(m_lock);
        m_start = now();
    }
    bool next_pixel(int prev_rays, int prev_samples, Pixel* pixel) noexcept override
    {
        std::lock_guard<std::mutex> lg(m_lock);
        m_ray_count += prev_rays;
        m_sample_count += prev_samples;
        if (m_current < m_pixels.size()) {
            *pixel = m_pixels[m_current++];
            ++m_pixel_count;
            return true;
        }
        return false;
    }
    bool next_pass() noexcept override
    {
        std::lock_guard<std::mutex> lg(m_lock);
        double t = now();
        m_prev = t - m_start;
        m_start = t;
        m_current = 0;
        if (m_pass_count == m_max_pass_count) return false;
        ++m_pass_count;
        return true;
    }
    void statistics(long long* pass_count, long long* pixel_count, long long* ray_count, long long* sample_count,
                    double* previous_pass_duration) const noexcept override
    {
        std::lock_guard<std::mutex> lg(m_lock);
        *pass_count = m_pass_count;
        *pixel_count = m_pixel_count;
        *ray_count = m_ray_count;
        *sample_count = m_sample_count;
        *previous_pass_duration = m_prev;
    }
};

// PainterNotifier (src/painter/painter.h:26-38)
template <size_t N>
struct PainterNotifier {
protected:
    virtual ~PainterNotifier() = default;

public:
    virtual void painter_pixel_before(const std::array<int_least16_t, N>& pixel) noexcept = 0;
    virtual void painter_pixel_after(const std::array<int_least16_t, N>& pixel, const Color& c) noexcept = 0;
    virtual void painter_error_message(const std::string& msg) noexcept = 0;
};

// PaintObjects (src/painter/objects.h:212-223)
template <size_t N, typename T>
struct PaintObjects {
    virtual ~PaintObjects() = default;
    virtual const std::vector<const GenericObject<N, T>*>& objects() const = 0;
    virtual const std::vector<const LightSource<N, T>*>& light_sources() const = 0;
    virtual const Projector<N, T>& projector() const = 0;
    virtual const SurfaceProperties<N, T>& default_surface_properties() const = 0;
};

namespace nsp_painter {

template <size_t N, typename T>
class SingleObjectScene final : public PaintObjects<N, T> {
    VisibleSharedMesh<N, T> m_object;
    std::unique_ptr<const Projector<N, T>> m_projector;
    std::unique_ptr<const LightSource<N, T>> m_light_source;
    SurfaceProperties<N, T> m_default;
    std::vector<const GenericObject<N, T>*> m_objects;
    std::vector<const LightSource<N, T>*> m_lights;

public:
    SingleObjectScene(const Color& background_color, const Color& default_color, float diffuse,
                      std::unique_ptr<const Projector<N, T>>&& projector,
                      std::unique_ptr<const LightSource<N, T>>&& light_source, std::shared_ptr<const Mesh<N, T>>&& mesh)
        : m_object(std::move(mesh)), m_projector(std::move(projector)), m_light_source(std::move(light_source))
    {
        // SingleObjectScene ctor, src/painter/scenes/single_object.h:64-84
        m_default.set_color(background_color);
        m_default.set_diffuse_and_fresnel(1, 0);
        m_default.set_light_source(true);
        m_default.set_light_source_color(Color(background_color.luminance()));
        m_object.set_color(default_color);
        m_object.set_diffuse_and_fresnel(diffuse, 0);
        m_object.set_light_source(false);
        m_lights.push_back(m_light_source.get());
        m_objects.push_back(&m_object);
    }
    const std::vector<const GenericObject<N, T>*>& objects() const override { return m_objects; }
    const std::vector<const LightSource<N, T>*>& light_sources() const override { return m_lights; }
    const Projector<N, T>& projector() const override { return *m_projector; }
    const SurfaceProperties<N, T>& default_surface_properties() const override { return m_default; }
};

}  // namespace nsp_painter

// single_object_scene with an explicit projector and light
// (src/painter/scenes/single_object.h:87-100)
template <size_t N, typename T>
std::unique_ptr<const PaintObjects<N, T>> single_object_scene(const Color& background_color, const Color& default_color,
                                                              float diffuse,
                                                              std::unique_ptr<const Projector<N, T>>&& projector,
                                                              std::unique_ptr<const LightSource<N, T>>&& light_source,
                                                              std::shared_ptr<const Mesh<N, T>> mesh)
{
    return std::make_unique<nsp_painter::SingleObjectScene<N, T>>(background_color, default_color, diffuse,
                                                                  std::move(projector), std::move(light_source),
                                                                  std::move(mesh));
}

// single_object_scene with automatic camera and light (single_object.h:102-179)
template <size_t N, typename T>
std::unique_ptr<const PaintObjects<N, T>> single_object_scene(const Color& background_color, const Color& default_color,
                                                              float diffuse, int min_screen_size, int max_screen_size,
                                                              std::shared_ptr<const Mesh<N, T>> mesh)
{
    if (min_screen_size < 3) throw ErrorException("Min screen size (" + std::to_string(min_screen_size) + ") is too small");
    if (min_screen_size > max_screen_size)
        throw ErrorException("Wrong min and max screen sizes: min = " + std::to_string(min_screen_size) +
                             ", max = " + std::to_string(max_screen_size));
    Vector<N, T> mn, mx;
    mesh->min_max(&mn, &mx);
    Vector<N, T> size, center;
    T len2 = 0;
    for (size_t i = 0; i < N; ++i) {
        size[i] = mx[i] - mn[i];
        center[i] = mn[i] + size[i] / T(2);
        len2 += size[i] * size[i];
    }
    T max_proj = -std::numeric_limits<T>::max();
    for (size_t i = 0; i + 1 < N; ++i) max_proj = std::max(size[i], max_proj);
    if (max_proj == 0) throw ErrorException("Object is a point on the screen");
    std::array<int, N - 1> screen;
    for (size_t i = 0; i + 1 < N; ++i) {
        int px = (int)std::lround((size[i] / max_proj) * max_screen_size);
        screen[i] = std::clamp(px, min_screen_size, max_screen_size);
    }
    Vector<N, T> cam = center;
    cam[N - 1] = mx[N - 1] + std::sqrt(len2);
    Vector<N, T> dir{};
    dir[N - 1] = -1;
    std::array<Vector<N, T>, N - 1> axes{};
    for (size_t i = 0; i + 1 < N; ++i) axes[i][i] = 1;
    T upp = max_proj / max_screen_size;
    Vector<N, T> light_pos;
    for (size_t i = 0; i < N; ++i) light_pos[i] = mx[i] + T(100) * (mx[i] - center[i]);
    return single_object_scene<N, T>(
        background_color, default_color, diffuse,
        std::make_unique<const VisibleParallelProjector<N, T>>(cam, dir, axes, upp, screen),
        std::make_unique<const VisibleConstantLight<N, T>>(light_pos, Color(1)), std::move(mesh));
}

// paint<N,T> (src/painter/painter.h:40-42, painter.cpp:599-625)
template <size_t N, typename T>
void paint(PainterNotifier<N - 1>* painter_notifier, int samples_per_pixel, const PaintObjects<N, T>& paint_objects,
           Paintbrush<N - 1>* paintbrush, int thread_count, std::atomic_bool* stop, bool smooth_normal) noexcept
{
    nsp_scene* scene = nullptr;
    try {
        if (!painter_notifier || !paintbrush || !stop) std::abort();  // ASSERT, painter.cpp:608
        if (thread_count < 1)
            throw ErrorException("Painter thread count (" + std::to_string(thread_count) + ") must be greater than 0");
        if (paintbrush->screen_size() != paint_objects.projector().screen_size())
            throw ErrorException("The paintbrush screen size are not equal to the projector screen size");
        if (paint_objects.objects().size() != 1)
            throw ErrorException("the GPU painter renders scenes of exactly one mesh object");
        const auto* obj = dynamic_cast<const VisibleSharedMesh<N, T>*>(paint_objects.objects()[0]);
        if (!obj) throw ErrorException("the GPU painter renders VisibleSharedMesh objects");
        std::vector<nsp_light> lights;
        for (const LightSource<N, T>* l : paint_objects.light_sources()) lights.push_back(l->data());
        nsp_projector pr = paint_objects.projector().data();
        nsp_surface os = obj->to_nsp();
        nsp_surface bg = paint_objects.default_surface_properties().to_nsp();
        nsp_painter::check(nsp_scene_create(nsp_painter::context(), (int)N, obj->mesh().handle(), &os,
                                            lights.data(), (int)lights.size(), &pr, &bg, &scene));
        const int64_t P = nsp_scene_pixel_count(scene);
        const std::array<int, N - 1> size = paintbrush->screen_size();
        std::vector<float> sum((size_t)P * 3);
        std::vector<int32_t> cnt((size_t)P);
        nsp_paint_params params{};
        params.samples_per_pixel = samples_per_pixel;
        params.max_depth = 100;  // MAX_RECURSION_LEVEL, painter.cpp:42
        params.smooth_normal = smooth_normal ? 1 : 0;
        params.precision = nsp_painter::fp64_flag() ? NSP_PRECISION_FP64 : NSP_PRECISION_FP32;
        params.passes = 1;
        params.seed = (uint64_t)std::chrono::steady_clock::now().time_since_epoch().count();
        std::atomic<int32_t> stop_i{0};
        paintbrush->first_pass();
        int64_t prev_rays = 0, prev_samples = 0;
        while (!*stop) {
            int64_t rays_before = 0, samples_before = 0;
            nsp_statistics(scene, nullptr, nullptr, &rays_before, &samples_before, nullptr);
            int rc = nsp_paint(scene, &params, nullptr, reinterpret_cast<const volatile int32_t*>(&stop_i));
            if (rc == NSP_E_STOPPED) break;
            nsp_painter::check(rc);
            int64_t rays_after = 0, samples_after = 0;
            nsp_statistics(scene, nullptr, nullptr, &rays_after, &samples_after, nullptr);
            nsp_painter::check(nsp_image_get(scene, sum.data(), cnt.data()));
            const int64_t pass_rays = rays_after - rays_before;
            const int64_t pixel_samples = (samples_after - samples_before) / P;
            std::array<int_least16_t, N - 1> pixel;
            int64_t delivered = 0;
            // deliver in the paintbrush's order; rays are attributed evenly per pixel
            while (!*stop && paintbrush->next_pixel((int)prev_rays, (int)prev_samples, &pixel)) {
                painter_notifier->painter_pixel_before(pixel);
                int64_t idx = 0, stride = 1;
                for (size_t i = 0; i + 1 < N; ++i) {
                    idx += (int64_t)pixel[i] * stride;
                    stride *= size[i];
                }
                float k = (float)std::max(1, cnt[idx]);
                painter_notifier->painter_pixel_after(pixel, Color(sum[idx * 3] / k, sum[idx * 3 + 1] / k,
                                                                   sum[idx * 3 + 2] / k));
                prev_rays = pass_rays / P + ((delivered < pass_rays % P) ? 1 : 0);
                prev_samples = pixel_samples;
                ++delivered;
            }
            if (*stop || !paintbrush->next_pass()) break;
        }
    } catch (const std::exception& e) {
        painter_notifier->painter_error_message(std::string("Painter error:\n") + e.what());
    } catch (...) {
        painter_notifier->painter_error_message("Unknown painter error");
    }
    nsp_scene_destroy(scene);
}
