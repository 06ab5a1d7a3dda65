// nsp_api.cpp -- C ABI of libnsp (include/nsp.h): contexts, meshes, scenes,
// the first-hit entry point and the progressive paint loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "nsp.h"
#include "nsp_build.hpp"
#include "nsp_internal.hpp"
#include "nsp_raster.hpp"
#include "nsp_render.hpp"

namespace nsp {

[[noreturn]] void raise(int code, const std::string& msg) { throw Error{code, msg}; }

void check_hip(hipError_t e, const char* what)
{
    if (e == hipSuccess) return;
    (void)hipGetLastError();
    int code = (e == hipErrorOutOfMemory) ? NSP_E_NOMEM : NSP_E_HIP;
    raise(code, std::string(what) + ": " + hipGetErrorString(e));
}

constexpr double PI_D = 3.14159265358979323846;

// TREE_MIN_OBJECTS_PER_BOX and tree_max_depth<N>(), src/painter/shapes/mesh.cpp:33-58
constexpr int TREE_MIN_OBJECTS_PER_BOX = 10;
static int tree_max_depth(int n)
{
    switch (n) {
    case 3: return 10;
    case 4: return 8;
    case 5: return 6;
    case 6: return 5;
    default: {
        double s = 1e9, r = std::pow(2.0, n);
        return (int)std::max(2.0, std::floor(std::log(s * (r - 1) + 1) / std::log(r)));
    }
    }
}

// cosine_sphere_coefficient, src/painter/coefficient/cosine_sphere.h:97-162
static long double cosine_sphere_coefficient(int n)
{
    long double p = (n & 1) == 0 ? 0.5L : 1.0L;
    for (int i = n - 1; i > 1; i -= 2) {
        p *= i;
        if (i > 2) p /= (i - 1);
    }
    if ((n & 1) == 0) p *= 3.14159265358979323846264338327950288L;
    return p;
}

// StratifiedJitteredSampler::one_dimension_size, src/painter/sampling/sampler.h:31-52
static int one_dimension_size(int dims, int spp)
{
    if (spp < 1) raise(NSP_E_INVALID, "Stratified jittered sample count (" + std::to_string(spp) + ") is not a positive integer");
    double v = std::pow(spp, 1.0 / dims);
    unsigned f = (unsigned)std::floor(v), c = (unsigned)std::ceil(v);
    unsigned long long p = 1;
    for (int i = 0; i < dims; ++i) p *= f;
    if (p >= (unsigned)spp) return (int)f;
    p = 1;
    for (int i = 0; i < dims; ++i) p *= c;
    if (p >= (unsigned)spp) return (int)c;
    raise(NSP_E_INVALID, "Could not compute one dimension sample count");
}

}  // namespace nsp

using namespace nsp;

// Device staging of nsp_trace_first_hit_noalloc (rays in, hits out) for at most
// NSP_NOALLOC_MAX_RAYS rays of up to NSP_MAX_DIM coordinates: org and dir
// (2 x 24 MiB), t (8 MiB), facet (4 MiB), near (1 MiB), each region 256-byte
// aligned.  One copy per device, held until the process ends; in host code the
// array's name is not a device address, so kernels and copies take the address
// hipGetSymbolAddress gives (nsp_context::scratch).
static constexpr int64_t NOALLOC_SCRATCH_BYTES = (int64_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

static constexpr int64_t align256(int64_t b) { return (b + 255) & ~(int64_t)255; }
static_assert(2 * align256((int64_t)sizeof(float) * NSP_NOALLOC_MAX_RAYS * NSP_MAX_DIM) +
                      align256((int64_t)sizeof(double) * NSP_NOALLOC_MAX_RAYS) +
                      align256((int64_t)sizeof(int32_t) * NSP_NOALLOC_MAX_RAYS) + align256(NSP_NOALLOC_MAX_RAYS) <=
                  NOALLOC_SCRATCH_BYTES,
              "the no-allocation staging must hold NSP_NOALLOC_MAX_RAYS rays");

struct nsp_context {
    int device = 0;
    unsigned char* scratch = nullptr;  // this device's copy of `scratch` (hipGetSymbolAddress)
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;  // the shadow stages of nsp_paint (joined back before each batch ends)
    std::string error;
    bool fail_next = false;
    // ray queues and per-path state of nsp_paint; every paint on this context
    // runs on its one stream, so all scenes share this workspace
    RenderBuffers work;
    // pinned staging of the per-batch pixel delivery (two slots) and its events
    HostBuf feed;
    hipEvent_t feed_ev[2] = {nullptr, nullptr};
};

struct nsp_mesh {
    nsp_context* ctx;
    int device = 0;
    MeshDev m;
};

struct nsp_scene {
    nsp_context* ctx;
    int device = 0;
    nsp_mesh* mesh;
    int n;
    nsp_surface object;
    std::vector<nsp_light> lights;
    nsp_projector proj;
    nsp_surface background;
    int64_t pixels = 0;
    DevBuf pix_sum, pix_cnt;
    DevBuf pix_aov;  // aov_channels(n) * pixels, allocated on the first pass with aov
    RasterPrep raster;  // rasterised primary visibility (boxes and work units), built on first use
    ShadowRasterPrep sraster;  // shadow rays rasterised from the light (footprints, grid), built on first use
    DevBuf pass_dev;  // int: the pass index the kernels read
    std::vector<ShapeRec> shapes;  // the objects after the mesh (nsp_scene_set_shapes)
    DevBuf shapes_dev;
    double shape_abs_max = 0;  // largest |coordinate| of the shapes' vertices
    DevBuf totals;  // 3 u64
    StageMarks marks;
    bool counters = false;
    // statistics (Paintbrush::statistics)
    int64_t pass_count = 0, pixel_count = 0, ray_count = 0, sample_count = 0;
    double last_pass_s = 0;
    nsp_stage_stats stage{};
};

// pass index of the counter-based sample streams of the scene's next pass
static int rng_pass(const nsp_scene* s) { return (int)s->pass_count; }

static thread_local std::string g_create_error;

namespace nsp {
// The device buffers of a scene's per-pixel sums (channel-major), for the
// denoiser's device-resident input: returns false when the scene has no AOV sums.
bool scene_pixel_buffers(const nsp_scene* s, ScenePixels* out)
{
    if (!s || !s->pix_aov.p) return false;
    out->ctx = s->ctx;
    out->n = s->n;
    out->pixels = s->pixels;
    for (int i = 0; i < s->n - 1; ++i) out->screen[i] = s->proj.screen_size[i];
    out->sum = s->pix_sum.as<float>();
    out->cnt = s->pix_cnt.as<int32_t>();
    out->aov = s->pix_aov.as<float>();
    return true;
}

hipStream_t context_stream(nsp_context* ctx) { return ctx->stream; }
int context_device(nsp_context* ctx) { return ctx->device; }
std::string* context_error(nsp_context* ctx) { return &ctx->error; }
}  // namespace nsp

// device workspace a default wavefront batch may take (bytes)
static constexpr int64_t DEFAULT_WORKSPACE = (int64_t)48 << 30;

template <class F>
static int guarded(nsp_context* ctx, F&& f)
{
    try {
        if (ctx) NSP_HIP(hipSetDevice(ctx->device));
        f();
        return NSP_OK;
    } catch (const Error& e) {
        if (ctx) ctx->error = e.msg;
        else g_create_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        if (ctx) ctx->error = "host allocation failed";
        return NSP_E_NOMEM;
    } catch (const std::exception& e) {
        if (ctx) ctx->error = e.what();
        return NSP_E_INVALID;
    } catch (...) {
        if (ctx) ctx->error = "unknown error";
        return NSP_E_INVALID;
    }
}

static void launch_gate(nsp_context* ctx)
{
    if (ctx->fail_next) {
        ctx->fail_next = false;
        raise(NSP_E_HIP, "kernel launch failed (injected)");
    }
}

template <typename T>
static void fill_scene(const nsp_scene& S, const nsp_paint_params& p, int pass, SceneK<T>& k)
{
    std::memset(&k, 0, sizeof k);
    const int n = S.n;
    const nsp_projector& pr = S.proj;
    double cd[NSP_MAX_DIM];
    {
        double r = pr.camera_dir[0] * pr.camera_dir[0];
        for (int i = 1; i < n; ++i) r = std::fma(pr.camera_dir[i], pr.camera_dir[i], r);
        double l = std::sqrt(r);
        for (int i = 0; i < n; ++i) cd[i] = pr.camera_dir[i] / l;
    }
    for (int i = 0; i < n; ++i) {
        k.cam_org[i] = (T)pr.camera_org[i];
        k.cam_dir[i] = (T)cd[i];
    }
    // projector constants (projector.h:110-136 perspective, 161-185 parallel, 212-249 spherical)
    k.proj = pr.kind;
    const double scale = pr.kind == NSP_PROJECTOR_PARALLEL ? pr.units_per_pixel : 1.0;
    for (int a = 0; a < n - 1; ++a) {
        double r = pr.screen_axes[a][0] * pr.screen_axes[a][0];
        for (int i = 1; i < n; ++i) r = std::fma(pr.screen_axes[a][i], pr.screen_axes[a][i], r);
        double l = std::sqrt(r);
        for (int i = 0; i < n; ++i) k.axes[a][i] = (T)((pr.screen_axes[a][i] / l) * scale);
        k.screen[a] = pr.screen_size[a];
    }
    const double half_angle = pr.view_angle_degrees * 0.5 * PI_D / 180;
    if (pr.kind == NSP_PROJECTOR_PERSPECTIVE) {
        const double dir_length = pr.screen_size[0] * 0.5 * std::tan(PI_D / 2 - half_angle);
        for (int i = 0; i < n; ++i) k.cam_dir[i] = (T)(cd[i] * dir_length);
    } else if (pr.kind == NSP_PROJECTOR_SPHERICAL) {
        const double r = pr.screen_size[0] * 0.5 / std::sin(half_angle);
        k.proj_r2 = (T)(r * r);
    }
    k.pixels = S.pixels;
    k.k = one_dimension_size(n - 1, p.samples_per_pixel);
    k.S = 1;
    for (int i = 0; i < n - 1; ++i) k.S *= k.k;
    k.k_inv = (T)1 / (T)k.k;
    k.light_count = (int)S.lights.size();
    for (int l = 0; l < k.light_count; ++l) {
        const nsp_light& L = S.lights[l];
        k.light_kind[l] = L.kind;
        for (int i = 0; i < n; ++i) k.light_pos[l][i] = (T)L.position[i];
        for (int c = 0; c < 3; ++c) k.light_rgb[l][c] = L.rgb[c];
        // PointLight: m_coef = pow(unit_intensity_distance, N - 1), light_source.h:37-40
        k.light_coef[l] = (T)std::pow(L.unit_intensity_distance, n - 1);
    }
    for (int c = 0; c < 3; ++c) {
        k.bg_rgb[c] = S.background.rgb[c];
        k.bg_light_rgb[c] = S.background.light_rgb[c];
        k.obj_rgb[c] = S.object.rgb[c];
        k.obj_light_rgb[c] = S.object.light_rgb[c];
    }
    k.bg_is_light = S.background.light_source;
    k.obj_is_light = S.object.light_source;
    k.diffuse = S.object.diffuse;
    // compute_ray_offset, painter.cpp:539-558: max |coordinate| of every object * 1000 * epsilon(T)
    double all_max = -INFINITY;
    for (int i = 0; i < n; ++i)
        all_max = std::max(all_max, std::max(std::fabs(S.mesh->m.mn[i]), std::fabs(S.mesh->m.mx[i])));
    if (!S.shapes.empty()) all_max = std::max(all_max, S.shape_abs_max);
    k.shapes = S.shapes.empty() ? nullptr : S.shapes_dev.as<ShapeRec>();
    k.shape_count = (int)S.shapes.size();
    const double eps = sizeof(T) == 4 ? 1.1920928955078125e-07 : 2.220446049250313e-16;
    k.ray_offset = (T)(all_max * (1000 * eps));
    k.cos_coef = (T)cosine_sphere_coefficient(n);
    k.max_depth = p.max_depth;
    k.smooth = p.smooth_normal ? 1 : 0;
    k.has_normals = S.mesh->m.has_normals ? 1 : 0;
    k.seed = p.seed;
    k.pass = pass;
}

extern "C" {

int nsp_context_create(int device, nsp_context** out)
{
    if (!out) return NSP_E_INVALID;
    *out = nullptr;
    return guarded(nullptr, [&] {
        int count = 0;
        NSP_HIP(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) raise(NSP_E_INVALID, "no such device: " + std::to_string(device));
        NSP_HIP(hipSetDevice(device));
        auto ctx = std::make_unique<nsp_context>();
        ctx->device = device;
        NSP_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        NSP_HIP(hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
        // the device address of this device's staging array, taken once here
        // with the device current
        void* sp = nullptr;
        NSP_HIP(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
        ctx->scratch = static_cast<unsigned char*>(sp);
        *out = ctx.release();
    });
}

void nsp_context_destroy(nsp_context* ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (hipEvent_t e : ctx->feed_ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* nsp_last_error(const nsp_context* ctx) { return ctx ? ctx->error.c_str() : g_create_error.c_str(); }

int nsp_debug_fail_next_launch(nsp_context* ctx)
{
    if (!ctx) return NSP_E_INVALID;
    ctx->fail_next = true;
    return NSP_OK;
}

int nsp_mesh_create(nsp_context* ctx, int n, const float* vertices, int64_t V, const int32_t* facets, int64_t F,
                    const float* normals, int64_t VN, const int32_t* facet_normals, const double* vertex_matrix,
                    nsp_mesh** out)
{
    if (!ctx || !out) return NSP_E_INVALID;
    *out = nullptr;
    return guarded(ctx, [&] {
        if (n < 3 || n > NSP_MAX_DIM) raise(NSP_E_INVALID, "dimension " + std::to_string(n) + " is not in [3, 6]");
        // create_mesh_object, src/painter/shapes/mesh.cpp:61-125
        if (V <= 0 || !vertices) raise(NSP_E_INVALID, "No vertices found in obj");
        if (F <= 0 || !facets) raise(NSP_E_INVALID, "No facets found in obj");
        if (F >= (int64_t)1 << 31) raise(NSP_E_INVALID, "too many facets");
        auto mesh = std::make_unique<nsp_mesh>();
        mesh->ctx = ctx;
        mesh->device = ctx->device;
        MeshDev& m = mesh->m;
        m.n = n;
        m.V = V;
        m.F = F;
        std::vector<double> vtx((size_t)V * n);
        if (vertex_matrix) {
            // MatrixMulVector, src/com/matrix_alg.h:30-63
            const double* M = vertex_matrix;
            if (M[n * (n + 1) + n] != 1) raise(NSP_E_INVALID, "Wrong matrix for matrix-vector multiplier");
            for (int i = 0; i < n; ++i)
                if (M[n * (n + 1) + i] != 0) raise(NSP_E_INVALID, "Wrong matrix for matrix-vector multiplier");
            for (int64_t v = 0; v < V; ++v)
                for (int i = 0; i < n; ++i) {
                    double r = M[i * (n + 1) + n];
                    for (int c = 0; c < n; ++c) r = std::fma(M[i * (n + 1) + c], (double)vertices[v * n + c], r);
                    vtx[v * n + i] = r;
                }
        } else {
            for (int64_t k = 0; k < V * n; ++k) vtx[k] = vertices[k];
        }
        for (int i = 0; i < n; ++i) {
            m.mn[i] = INFINITY;
            m.mx[i] = -INFINITY;
        }
        for (int64_t f = 0; f < F; ++f)
            for (int k = 0; k < n; ++k) {
                int32_t vi = facets[f * n + k];
                if (vi < 0 || vi >= V) raise(NSP_E_INVALID, "facet vertex index out of range");
                for (int i = 0; i < n; ++i) {
                    m.mn[i] = std::min(m.mn[i], vtx[(size_t)vi * n + i]);
                    m.mx[i] = std::max(m.mx[i], vtx[(size_t)vi * n + i]);
                }
            }
        hipStream_t s = ctx->stream;
        m.vtx64.alloc(sizeof(double) * V * n);
        NSP_HIP(hipMemcpyAsync(m.vtx64.p, vtx.data(), sizeof(double) * V * n, hipMemcpyHostToDevice, s));
        m.fac.alloc(sizeof(int32_t) * F * n);
        NSP_HIP(hipMemcpyAsync(m.fac.p, facets, sizeof(int32_t) * F * n, hipMemcpyHostToDevice, s));
        std::vector<double> vn;
        if (normals && facet_normals && VN > 0) {
            // -1 marks a facet vertex without a normal (such facets are shaded flat)
            for (int64_t k = 0; k < F * n; ++k)
                if (facet_normals[k] < -1 || facet_normals[k] >= VN)
                    raise(NSP_E_INVALID, "facet normal index out of range");
            m.has_normals = true;
            m.VN = VN;
            vn.resize((size_t)VN * n);
            for (int64_t k = 0; k < VN * n; ++k) vn[k] = normals[k];
            m.vn64.alloc(sizeof(double) * VN * n);
            NSP_HIP(hipMemcpyAsync(m.vn64.p, vn.data(), sizeof(double) * VN * n, hipMemcpyHostToDevice, s));
            m.fnrm.alloc(sizeof(int32_t) * F * n);
            NSP_HIP(hipMemcpyAsync(m.fnrm.p, facet_normals, sizeof(int32_t) * F * n, hipMemcpyHostToDevice, s));
        }
        NSP_HIP(hipStreamSynchronize(s));
        launch_gate(ctx);
        build_mesh_device(m, tree_max_depth(n), TREE_MIN_OBJECTS_PER_BOX, s);
        *out = mesh.release();
    });
}

void nsp_mesh_destroy(nsp_mesh* mesh)
{
    if (!mesh) return;
    // no use of mesh->ctx here: the context may already be gone
    (void)hipSetDevice(mesh->device);
    (void)hipDeviceSynchronize();
    delete mesh;
}

int nsp_mesh_min_max(const nsp_mesh* mesh, double* mn, double* mx)
{
    if (!mesh || !mn || !mx) return NSP_E_INVALID;
    for (int i = 0; i < mesh->m.n; ++i) {
        mn[i] = mesh->m.mn[i];
        mx[i] = mesh->m.mx[i];
    }
    return NSP_OK;
}

int nsp_mesh_get_info(const nsp_mesh* mesh, nsp_mesh_info* info)
{
    if (!mesh || !info) return NSP_E_INVALID;
    const MeshDev& m = mesh->m;
    std::memset(info, 0, sizeof *info);
    info->n = m.n;
    info->facets = m.F;
    info->tree_nodes = m.tree.node_count;
    info->tree_leaves = m.tree.leaf_count;
    info->tree_entries = m.tree.entry_count;
    info->tree_levels = m.tree.levels;
    info->basis_ms = m.basis_ms;
    info->tree_ms = m.build_ms;
    for (int i = 0; i < m.n; ++i) {
        info->root_lo[i] = m.tree.root_lo[i];
        info->root_size[i] = m.tree.root_size[i];
    }
    return NSP_OK;
}

int nsp_mesh_set_record_layout(nsp_mesh* mesh, int layout)
{
    if (!mesh) return NSP_E_INVALID;
    if (layout != NSP_LAYOUT_LEAF_SOA && layout != NSP_LAYOUT_SIMPLEX_AOS) {
        if (mesh->ctx) mesh->ctx->error = "unknown record layout " + std::to_string(layout);
        return NSP_E_INVALID;
    }
    return guarded(mesh->ctx, [&] {
        if (layout == NSP_LAYOUT_LEAF_SOA) build_soa_device(mesh->m, mesh->ctx->stream);
        mesh->m.layout = layout;
    });
}

int nsp_mesh_get_record_layout(const nsp_mesh* mesh) { return mesh ? mesh->m.layout : NSP_E_INVALID; }

int nsp_mesh_records(const nsp_mesh* mesh, double* normal, double* plane_n, double* plane_d)
{
    if (!mesh) return NSP_E_INVALID;
    nsp_context* ctx = mesh->ctx;
    return guarded(ctx, [&] {
        const MeshDev& m = mesh->m;
        const int n = m.n, st = rec64_stride(n);
        std::vector<double> rec((size_t)m.F * st);
        NSP_HIP(hipMemcpy(rec.data(), m.rec64.p, sizeof(double) * rec.size(), hipMemcpyDeviceToHost));
        for (int64_t f = 0; f < m.F; ++f) {
            const double* r = rec.data() + f * st;
            if (normal) std::memcpy(normal + f * n, r, sizeof(double) * n);
            if (plane_n) std::memcpy(plane_n + f * (n - 1) * n, r + 2 * n, sizeof(double) * (n - 1) * n);
            if (plane_d) std::memcpy(plane_d + f * (n - 1), r + 2 * n + (n - 1) * n, sizeof(double) * (n - 1));
        }
    });
}

int64_t nsp_mesh_tree_dump(const nsp_mesh* mesh, int64_t* out, int64_t cap)
{
    if (!mesh) return NSP_E_INVALID;
    int64_t written = 0;
    int rc = guarded(mesh->ctx, [&] {
        const MeshDev& m = mesh->m;
        const TreeDev& t = m.tree;
        const int n = m.n;
        std::vector<int2> nodes(t.node_count);
        std::vector<int8_t> lv(t.node_count);
        std::vector<int32_t> q((size_t)t.node_count * n);
        std::vector<int32_t> ent(std::max<int64_t>(t.entry_count, 1));
        NSP_HIP(hipMemcpy(nodes.data(), t.nodes.p, sizeof(int2) * t.node_count, hipMemcpyDeviceToHost));
        NSP_HIP(hipMemcpy(lv.data(), t.node_level.p, t.node_count, hipMemcpyDeviceToHost));
        NSP_HIP(hipMemcpy(q.data(), t.node_q.p, sizeof(int32_t) * q.size(), hipMemcpyDeviceToHost));
        if (t.entry_count)
            NSP_HIP(hipMemcpy(ent.data(), t.entries.p, sizeof(int32_t) * t.entry_count, hipMemcpyDeviceToHost));
        // leaves in depth-first child order (the order of the reference's recursive build)
        std::vector<int64_t> stack{0};
        while (!stack.empty()) {
            int64_t id = stack.back();
            stack.pop_back();
            int2 nd = nodes[id];
            if (nd.y < 0) {
                int children = 1 << n;
                for (int c = children - 1; c >= 0; --c) stack.push_back((int64_t)nd.x + (int64_t)c * (-nd.y));
                continue;
            }
            int64_t need = 2 + n + nd.y;
            if (out && written + need <= cap) {
                out[written] = lv[id];
                for (int k = 0; k < n; ++k) out[written + 1 + k] = q[(size_t)id * n + k];
                out[written + 1 + n] = nd.y;
                for (int j = 0; j < nd.y; ++j) out[written + 2 + n + j] = ent[nd.x + j];
            }
            written += need;
        }
    });
    return rc == NSP_OK ? written : rc;
}

int nsp_trace_first_hit_device(nsp_mesh* mesh, const float* org, const float* dir, int64_t count, int32_t* facet,
                               double* t, uint8_t* near, void* stream)
{
    if (!mesh || count < 0) return NSP_E_INVALID;
    if (count == 0) return NSP_OK;
    if (!org || !dir || !facet || !t) return NSP_E_INVALID;
    nsp_context* ctx = mesh->ctx;
    return guarded(ctx, [&] {
        launch_gate(ctx);
        hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
        launch_first_hit(mesh->m, org, dir, count, facet, t, near, nullptr, s);
    });
}

int nsp_trace_first_hit(nsp_mesh* mesh, const float* org, const float* dir, int64_t count, int32_t* facet, double* t,
                        uint8_t* near)
{
    if (!mesh || count < 0) return NSP_E_INVALID;
    if (count == 0) return NSP_OK;
    if (!org || !dir || !facet || !t) return NSP_E_INVALID;
    nsp_context* ctx = mesh->ctx;
    return guarded(ctx, [&] {
        launch_gate(ctx);
        const int n = mesh->m.n;
        hipStream_t s = ctx->stream;
        DevBuf d_o, d_d, d_i, d_t, d_n;
        d_o.alloc(sizeof(float) * count * n);
        d_d.alloc(sizeof(float) * count * n);
        d_i.alloc(sizeof(int32_t) * count);
        d_t.alloc(sizeof(double) * count);
        d_n.alloc(count);
        NSP_HIP(hipMemcpyAsync(d_o.p, org, sizeof(float) * count * n, hipMemcpyHostToDevice, s));
        NSP_HIP(hipMemcpyAsync(d_d.p, dir, sizeof(float) * count * n, hipMemcpyHostToDevice, s));
        launch_first_hit(mesh->m, d_o.as<float>(), d_d.as<float>(), count, d_i.as<int32_t>(), d_t.as<double>(),
                         d_n.as<uint8_t>(), nullptr, s);
        NSP_HIP(hipMemcpyAsync(facet, d_i.p, sizeof(int32_t) * count, hipMemcpyDeviceToHost, s));
        NSP_HIP(hipMemcpyAsync(t, d_t.p, sizeof(double) * count, hipMemcpyDeviceToHost, s));
        if (near) NSP_HIP(hipMemcpyAsync(near, d_n.p, count, hipMemcpyDeviceToHost, s));
        NSP_HIP(hipStreamSynchronize(s));
    });
}

int nsp_trace_first_hit_noalloc(nsp_mesh* mesh, const float* org, const float* dir, int64_t count, int32_t* facet,
                                double* t, uint8_t* near)
{
    if (!mesh || count < 0) return NSP_E_INVALID;
    if (count == 0) return NSP_OK;
    if (!org || !dir || !facet || !t) return NSP_E_INVALID;
    nsp_context* ctx = mesh->ctx;
    if (count > NSP_NOALLOC_MAX_RAYS) {
        ctx->error = "nsp_trace_first_hit_noalloc takes at most " + std::to_string(NSP_NOALLOC_MAX_RAYS) +
                     " rays per call (" + std::to_string(count) + " given)";
        return NSP_E_INVALID;
    }
    return guarded(ctx, [&] {
        launch_gate(ctx);
        if (!ctx->scratch) raise(NSP_E_HIP, "the device staging array has no address");
        const int n = mesh->m.n;
        hipStream_t s = ctx->stream;
        // regions of the staging array sized for this call; the kernel writes
        // every facet, t and near entry below count, so nothing is cleared
        unsigned char* base = ctx->scratch;
        int64_t off = 0;
        float* d_o = reinterpret_cast<float*>(base + off);
        off += align256((int64_t)sizeof(float) * count * n);
        float* d_d = reinterpret_cast<float*>(base + off);
        off += align256((int64_t)sizeof(float) * count * n);
        double* d_t = reinterpret_cast<double*>(base + off);
        off += align256((int64_t)sizeof(double) * count);
        int32_t* d_i = reinterpret_cast<int32_t*>(base + off);
        off += align256((int64_t)sizeof(int32_t) * count);
        uint8_t* d_n = base + off;
        NSP_HIP(hipMemcpyAsync(d_o, org, sizeof(float) * count * n, hipMemcpyHostToDevice, s));
        NSP_HIP(hipMemcpyAsync(d_d, dir, sizeof(float) * count * n, hipMemcpyHostToDevice, s));
        launch_first_hit(mesh->m, d_o, d_d, count, d_i, d_t, d_n, nullptr, s);
        NSP_HIP(hipMemcpyAsync(facet, d_i, sizeof(int32_t) * count, hipMemcpyDeviceToHost, s));
        NSP_HIP(hipMemcpyAsync(t, d_t, sizeof(double) * count, hipMemcpyDeviceToHost, s));
        if (near) NSP_HIP(hipMemcpyAsync(near, d_n, count, hipMemcpyDeviceToHost, s));
        NSP_HIP(hipStreamSynchronize(s));
    });
}

int nsp_scene_create(nsp_context* ctx, int n, nsp_mesh* mesh, const nsp_surface* object, const nsp_light* lights,
                     int light_count, const nsp_projector* projector, const nsp_surface* default_surface,
                     nsp_scene** out)
{
    if (!ctx || !out) return NSP_E_INVALID;
    *out = nullptr;
    return guarded(ctx, [&] {
        if (!mesh || !object || !projector || !default_surface) raise(NSP_E_INVALID, "null scene argument");
        if (mesh->m.n != n) raise(NSP_E_INVALID, "mesh dimension differs from scene dimension");
        if (light_count < 0 || (light_count > 0 && !lights)) raise(NSP_E_INVALID, "bad light list");
        if (light_count > MAX_LIGHTS)
            raise(NSP_E_UNSUPPORTED, "at most " + std::to_string(MAX_LIGHTS) + " light sources are supported");
        for (int l = 0; l < light_count; ++l)
            if (lights[l].kind != NSP_LIGHT_POINT && lights[l].kind != NSP_LIGHT_CONSTANT)
                raise(NSP_E_INVALID, "unknown light kind " + std::to_string(lights[l].kind));
        if (projector->kind != NSP_PROJECTOR_PARALLEL && projector->kind != NSP_PROJECTOR_PERSPECTIVE &&
            projector->kind != NSP_PROJECTOR_SPHERICAL)
            raise(NSP_E_INVALID, "unknown projector kind " + std::to_string(projector->kind));
        int64_t pixels = 1;
        for (int i = 0; i < n - 1; ++i) {
            // screen_org, src/painter/projectors/projector.h:51-66; pixel coordinates are int_least16_t
            if (projector->screen_size[i] < 1)
                raise(NSP_E_INVALID, "Projection size " + std::to_string(i) + " is not positive (" +
                                         std::to_string(projector->screen_size[i]) + ")");
            if (projector->screen_size[i] > 32767) raise(NSP_E_INVALID, "screen size exceeds int_least16_t pixels");
            pixels *= projector->screen_size[i];
        }
        const double angle = projector->view_angle_degrees;
        if (projector->kind == NSP_PROJECTOR_PARALLEL) {
            if (!(projector->units_per_pixel > 0))
                raise(NSP_E_INVALID, "Error units per pixel for parallel projection");
        } else if (projector->kind == NSP_PROJECTOR_PERSPECTIVE) {
            // PerspectiveProjector, projector.h:120-136
            if (!(angle > 0 && angle < 180))
                raise(NSP_E_INVALID, "Perspective projection: error view angle " + std::to_string(angle));
        } else {
            // SphericalProjector, projector.h:222-249
            const double sin_alpha = std::sin(angle * 0.5 * PI_D / 180);
            if (!(angle > 0))
                raise(NSP_E_INVALID, "Spherical projection view angle " + std::to_string(angle) + " is not positive");
            const double k = sin_alpha / projector->screen_size[0];
            double r = sin_alpha * sin_alpha;
            for (int i = 1; i < n - 1; ++i) r += (k * projector->screen_size[i]) * (k * projector->screen_size[i]);
            if (!(r < 1))
                raise(NSP_E_INVALID, "Spherical projection view angle " + std::to_string(angle) + " is too big");
        }
        // unit_dir_and_axes + check_vectors_orthogonal, projector.h:30-86
        auto len = [n](const double* v) {
            double r = v[0] * v[0];
            for (int i = 1; i < n; ++i) r = std::fma(v[i], v[i], r);
            return std::sqrt(r);
        };
        double cd[NSP_MAX_DIM], ax[NSP_MAX_DIM - 1][NSP_MAX_DIM];
        double l0 = len(projector->camera_dir);
        for (int i = 0; i < n; ++i) cd[i] = projector->camera_dir[i] / l0;
        for (int a = 0; a < n - 1; ++a) {
            double la = len(projector->screen_axes[a]);
            for (int i = 0; i < n; ++i) ax[a][i] = projector->screen_axes[a][i] / la;
        }
        const double LIMIT_COS = 2.220446049250313e-16 * 100;
        auto dotn = [n](const double* a, const double* b) {
            double r = a[0] * b[0];
            for (int i = 1; i < n; ++i) r = std::fma(a[i], b[i], r);
            return r;
        };
        for (int a = 0; a < n - 1; ++a) {
            if (!(std::fabs(dotn(ax[a], cd)) <= LIMIT_COS))
                raise(NSP_E_INVALID, "The screen axis " + std::to_string(a) + " is not orthogonal to the camera direction");
            for (int b = a + 1; b < n - 1; ++b)
                if (!(std::fabs(dotn(ax[a], ax[b])) <= LIMIT_COS))
                    raise(NSP_E_INVALID, "The screen axis " + std::to_string(a) + " is not orthogonal to the screen axes " +
                                             std::to_string(b));
        }
        auto sc = std::make_unique<nsp_scene>();
        sc->ctx = ctx;
        sc->device = ctx->device;
        sc->mesh = mesh;
        sc->n = n;
        sc->object = *object;
        sc->lights.assign(lights, lights + light_count);
        sc->proj = *projector;
        sc->background = *default_surface;
        sc->pixels = pixels;
        sc->pix_sum.alloc(sizeof(float) * 3 * pixels);
        sc->pix_cnt.alloc(sizeof(int32_t) * pixels);
        NSP_HIP(hipMemsetAsync(sc->pix_sum.p, 0, sizeof(float) * 3 * pixels, ctx->stream));
        NSP_HIP(hipMemsetAsync(sc->pix_cnt.p, 0, sizeof(int32_t) * pixels, ctx->stream));
        sc->totals.alloc(sizeof(unsigned long long) * 3);
        NSP_HIP(hipStreamSynchronize(ctx->stream));
        *out = sc.release();
    });
}

}  // extern "C"

// Parallelotope::create_planes (parallelotope.h:129-153) and
// HyperplaneParallelotope::set_data with HyperplaneSimplexGeometry::set_data
// (hyperplane_parallelotope.h:68-74, hyperplane_geometry.h:80-110), FP64.
template <int N>
static ShapeRec make_shape(const nsp_shape& in)
{
    ShapeRec r{};
    r.kind = in.kind;
    r.light = in.surface.light_source != 0;
    for (int c = 0; c < 3; ++c) {
        r.rgb[c] = in.surface.rgb[c];
        r.light_rgb[c] = in.surface.light_rgb[c];
    }
    r.diffuse = in.surface.diffuse;
    for (int k = 0; k < N; ++k) r.org[k] = in.org[k];
    if (in.kind == NSP_SHAPE_PARALLELOTOPE) {
        for (int i = 0; i < N; ++i) {
            double m[(N - 1) * N];
            for (int a = 0, row = 0; a < N; ++a)
                if (a != i) {
                    for (int k = 0; k < N; ++k) m[row * N + k] = in.edges[a][k];
                    ++row;
                }
            double nn[N];
            ortho_nn<N, double>(m, nn);
            normalize<N, double>(nn);
            double ov[N];
            for (int k = 0; k < N; ++k) ov[k] = r.org[k] + in.edges[i][k];
            double d1 = dot<N, double>(r.org, nn);
            double d2 = -dot<N, double>(ov, nn);
            if (dot<N, double>(nn, in.edges[i]) > 0) {
                for (int k = 0; k < N; ++k) nn[k] = -nn[k];
                d1 = -d1;
                d2 = -d2;
            }
            for (int k = 0; k < N; ++k) r.pn[i][k] = nn[k];
            r.d1[i] = d1;
            r.d2[i] = d2;
        }
    } else {
        double vec[(N - 1) * N];
        for (int a = 0; a < N - 1; ++a)
            for (int k = 0; k < N; ++k) vec[a * N + k] = in.edges[a][k];
        ortho_nn<N, double>(vec, r.n);
        normalize<N, double>(r.n);
        // the simplex geometry of the points org + e_i and org
        double pts[N][N];
        for (int a = 0; a < N - 1; ++a)
            for (int k = 0; k < N; ++k) pts[a][k] = r.org[k] + in.edges[a][k];
        for (int k = 0; k < N; ++k) pts[N - 1][k] = r.org[k];
        for (int a = 0; a < N - 1; ++a)
            for (int k = 0; k < N; ++k) vec[a * N + k] = pts[a][k] - pts[N - 1][k];
        for (int i = 0; i < N - 1; ++i) {
            double m[(N - 1) * N];
            for (int a = 0; a < N - 1; ++a)
                for (int k = 0; k < N; ++k) m[a * N + k] = a == i ? r.n[k] : vec[a * N + k];
            double pn[N];
            ortho_nn<N, double>(m, pn);
            double d = dot<N, double>(pts[N - 1], pn);
            const double dist = dot<N, double>(pts[i], pn) - d;
            for (int k = 0; k < N; ++k) r.pn[i][k] = pn[k] / dist;
            r.d1[i] = d / dist;
        }
    }
    return r;
}

extern "C" {

int nsp_scene_set_shapes(nsp_scene* scene, const nsp_shape* shapes, int count)
{
    if (!scene) return NSP_E_INVALID;
    return guarded(scene->ctx, [&] {
        if (count < 0 || count > NSP_MAX_SHAPES || (count > 0 && !shapes))
            raise(NSP_E_INVALID, "shape count must be in [0, " + std::to_string(NSP_MAX_SHAPES) + "]");
        const int n = scene->n;
        std::vector<ShapeRec> recs;
        double abs_max = 0;
        for (int j = 0; j < count; ++j) {
            const nsp_shape& sh = shapes[j];
            if (sh.kind != NSP_SHAPE_PARALLELOTOPE && sh.kind != NSP_SHAPE_HYPERPLANE_PARALLELOTOPE)
                raise(NSP_E_INVALID, "unknown shape kind " + std::to_string(sh.kind));
            ShapeRec r{};
            switch (n) {
            case 3: r = make_shape<3>(sh); break;
            case 4: r = make_shape<4>(sh); break;
            case 5: r = make_shape<5>(sh); break;
            case 6: r = make_shape<6>(sh); break;
            default: raise(NSP_E_INVALID, "dimension not supported");
            }
            bool finite = true;
            for (int a = 0; a < n; ++a) {
                finite = finite && std::isfinite(r.n[a]) && std::isfinite(r.d1[a]) && std::isfinite(r.d2[a]);
                for (int k = 0; k < n; ++k) finite = finite && std::isfinite(r.pn[a][k]);
            }
            if (!finite) raise(NSP_E_INVALID, "shape " + std::to_string(j) + " has dependent edge vectors");
            // the vertices: org plus every subset of the edges (parallelotope_vertices,
            // hyperplane_parallelotope_vertices), for compute_ray_offset
            const int edges = sh.kind == NSP_SHAPE_PARALLELOTOPE ? n : n - 1;
            for (int mask = 0; mask < (1 << edges); ++mask)
                for (int k = 0; k < n; ++k) {
                    double x = sh.org[k];
                    for (int a = 0; a < edges; ++a)
                        if (mask & (1 << a)) x += sh.edges[a][k];
                    abs_max = std::max(abs_max, std::fabs(x));
                }
            recs.push_back(r);
        }
        hipStream_t s = scene->ctx->stream;
        NSP_HIP(hipStreamSynchronize(s));
        if (!recs.empty()) {
            scene->shapes_dev.alloc(sizeof(ShapeRec) * recs.size());
            NSP_HIP(hipMemcpyAsync(scene->shapes_dev.p, recs.data(), sizeof(ShapeRec) * recs.size(),
                                   hipMemcpyHostToDevice, s));
            NSP_HIP(hipStreamSynchronize(s));
        }
        scene->shapes = std::move(recs);
        scene->shape_abs_max = abs_max;
    });
}

void nsp_scene_destroy(nsp_scene* scene)
{
    if (!scene) return;
    (void)hipSetDevice(scene->device);
    (void)hipDeviceSynchronize();
    delete scene;
}

int64_t nsp_scene_pixel_count(const nsp_scene* scene) { return scene ? scene->pixels : NSP_E_INVALID; }

// Delivery of the pixels a wavefront batch finished (PainterNotifier::painter_pixel_after
// for each pixel, src/painter/painter.cpp:410-427): after a batch's launches its
// pixel range is copied into one of two pinned slots and an event is recorded;
// the previous batch's slot is delivered while this batch runs.
namespace {
struct PixelFeed {
    nsp_context* ctx;
    const nsp_scene* scene;
    const nsp_callbacks* cb;
    int64_t batch_pixels;
    int64_t pending_first = -1, pending_count = 0;
    int pending_slot = 0;
    std::vector<float> mean;

    PixelFeed(nsp_context* c, const nsp_scene* sc, const nsp_callbacks* k, int64_t bp)
        : ctx(c), scene(sc), cb(k), batch_pixels(bp)
    {
        ctx->feed.alloc((size_t)2 * bp * (3 * sizeof(float) + sizeof(int32_t)));
        for (hipEvent_t& e : ctx->feed_ev)
            if (!e) NSP_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    float* slot_sum(int k) const { return ctx->feed.as<float>() + (size_t)k * batch_pixels * 4; }
    int32_t* slot_cnt(int k) const { return reinterpret_cast<int32_t*>(slot_sum(k) + 3 * batch_pixels); }

    // queue the copies of pixels [first, first + count) after the work on s
    void batch_done(int64_t first, int64_t count, hipStream_t s)
    {
        const int k = pending_first < 0 ? 0 : 1 - pending_slot;
        const int64_t P = scene->pixels;
        for (int c = 0; c < 3; ++c)
            NSP_HIP(hipMemcpyAsync(slot_sum(k) + c * batch_pixels, scene->pix_sum.as<float>() + c * P + first,
                                   sizeof(float) * count, hipMemcpyDeviceToHost, s));
        NSP_HIP(hipMemcpyAsync(slot_cnt(k), scene->pix_cnt.as<int32_t>() + first, sizeof(int32_t) * count,
                               hipMemcpyDeviceToHost, s));
        NSP_HIP(hipEventRecord(ctx->feed_ev[k], s));
        deliver();
        pending_first = first;
        pending_count = count;
        pending_slot = k;
    }
    // hand the pending slot to the callback, in chunks of at most 65536 pixels
    void deliver()
    {
        if (pending_first < 0) return;
        NSP_HIP(hipEventSynchronize(ctx->feed_ev[pending_slot]));
        const float* sum = slot_sum(pending_slot);
        const int32_t* cnt = slot_cnt(pending_slot);
        const int64_t CH = 65536;
        mean.resize((size_t)3 * std::min(CH, pending_count));
        for (int64_t b = 0; b < pending_count; b += CH) {
            const int64_t c = std::min(CH, pending_count - b);
            for (int64_t j = 0; j < c; ++j) {
                const int32_t k = std::max(1, cnt[b + j]);
                for (int ch = 0; ch < 3; ++ch) mean[(size_t)(j * 3 + ch)] = sum[ch * batch_pixels + b + j] / k;
            }
            cb->pixels_after(cb->user, pending_first + b, c, mean.data());
        }
        pending_first = -1;
    }
};

struct GraphHolder {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~GraphHolder()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};
}  // namespace

int nsp_paint(nsp_scene* scene, const nsp_paint_params* params, const nsp_callbacks* cb, const volatile int32_t* stop)
{
    if (!scene || !params) return NSP_E_INVALID;
    nsp_context* ctx = scene->ctx;
    int rc = guarded(ctx, [&] {
        const nsp_paint_params& p = *params;
        if (p.max_depth < 0 || p.max_depth > 250) raise(NSP_E_INVALID, "max_depth must be in [0, 250]");
        // BarPaintbrush max_pass_count: -1 or positive (src/painter/paintbrushes/paintbrush.h:141-144)
        if (!(p.passes == -1 || p.passes > 0)) raise(NSP_E_INVALID, "Error paintbrush max pass count " + std::to_string(p.passes));
        if (p.passes == -1 && !stop) raise(NSP_E_INVALID, "passes = -1 paints until *stop is set and needs a stop flag");
        if (p.precision != NSP_PRECISION_FP32 && p.precision != NSP_PRECISION_FP64)
            raise(NSP_E_INVALID, "unknown precision");
        if (p.primary != NSP_PRIMARY_AUTO && p.primary != NSP_PRIMARY_WALK) raise(NSP_E_INVALID, "unknown primary mode");
        if (!scene->shapes.empty() && p.smooth_normal && scene->mesh->m.has_normals)
            raise(NSP_E_UNSUPPORTED, "smooth-normal renders of meshes with vertex normals do not support shapes");
        launch_gate(ctx);
        const int n = scene->n;
        hipStream_t s = ctx->stream;
        if (!scene->pass_dev.p) scene->pass_dev.alloc(sizeof(int));
        SceneK<float> kf;
        SceneK<double> kd;
        fill_scene<float>(*scene, p, 0, kf);
        fill_scene<double>(*scene, p, 0, kd);
        kf.pass_dev = kd.pass_dev = scene->pass_dev.as<int>();
        const int S = kf.S;
        const bool raster = p.precision == NSP_PRECISION_FP32 && p.primary == NSP_PRIMARY_AUTO &&
                            scene->proj.kind == NSP_PROJECTOR_PARALLEL;
        // shadow rays rasterised from the light: FP32 walks over simplex-ordered
        // records, one point light, no shapes, no smooth-normal probes (their
        // rays do not start at the hit; the rasteriser rebuilds each ray from
        // its hit point), and the light sees the mesh in front of it, decided
        // by shadow_raster_prepare
        bool sraster = p.precision == NSP_PRECISION_FP32 && p.primary == NSP_PRIMARY_AUTO &&
                       scene->lights.size() == 1 && scene->lights[0].kind == NSP_LIGHT_POINT &&
                       scene->shapes.empty() && scene->mesh->m.layout == NSP_LAYOUT_SIMPLEX_AOS &&
                       !(p.smooth_normal && scene->mesh->m.has_normals);
        // Default batch: the largest power of two of paths, at most 2^27 (FP32) or
        // 2^26 (FP64), whose workspace fits DEFAULT_WORKSPACE.  Large batches keep
        // the persistent walks busy and leave fewer batch tails (C1, 27 samples per
        // pixel: 898 Mrays/s at 2^23 paths, 1076 at 2^24, 1218 at 2^25, 1308 at 2^26;
        // C2 in one 2^27-path batch 282 against 276 in two of 2^26).
        int64_t want = p.batch_paths;
        if (want <= 0) {
            const int64_t per =
                render_bytes_per_path(n, p.precision, (int)scene->lights.size(), p.aov != 0, raster, sraster);
            want = (int64_t)1 << (p.precision ? 26 : 27);
            while (want > ((int64_t)1 << 20) && want * per > DEFAULT_WORKSPACE) want >>= 1;
        }
        int64_t batch_pixels = std::max<int64_t>(1, want / S);
        batch_pixels = std::min<int64_t>(batch_pixels, scene->pixels);
        const int64_t cap = batch_pixels * S;
        alloc_render_buffers(ctx->work, n, p.precision, cap, p.max_depth, (int)scene->lights.size());
        if (raster) {
            raster_prepare(scene->raster, scene->mesh->m, kd, S, s);
            ctx->work.keys.alloc(sizeof(unsigned long long) * cap);
        }
        if (sraster) sraster = shadow_raster_prepare(scene->sraster, scene->mesh->m, kd.light_pos[0], s);
        if (sraster) shadow_raster_buffers(ctx->work, n, cap);
        if (p.aov) {
            ctx->work.aov.alloc(sizeof(float) * aov_channels(n) * ctx->work.cap);  // per-path planes of stride cap
            if (!scene->pix_aov.p) {
                scene->pix_aov.alloc(sizeof(float) * pix_aov_channels(n) * scene->pixels);
                NSP_HIP(hipMemsetAsync(scene->pix_aov.p, 0, sizeof(float) * pix_aov_channels(n) * scene->pixels, s));
            }
        }
        // the batches live in device memory; each batch's launch sequence
        // selects its batch there and the kernels read the pass index there,
        // so one launch sequence (and one captured graph) serves every batch of
        // every pass
        std::vector<BatchDesc> descs;
        for (int64_t px = 0; px < scene->pixels; px += batch_pixels)
            descs.push_back(BatchDesc{px, std::min(batch_pixels, scene->pixels - px) * S});
        ctx->work.batch_list.alloc(sizeof(BatchDesc) * descs.size());
        NSP_HIP(hipMemcpyAsync(ctx->work.batch_list.p, descs.data(), sizeof(BatchDesc) * descs.size(),
                               hipMemcpyHostToDevice, s));
        RenderLaunch rl;
        rl.mesh = &scene->mesh->m;
        rl.scene = p.precision ? (const void*)&kd : (const void*)&kf;
        rl.buf = &ctx->work;
        rl.pix_sum = scene->pix_sum.as<float>();
        rl.pix_cnt = scene->pix_cnt.as<int32_t>();
        rl.pix_aov = p.aov ? scene->pix_aov.as<float>() : nullptr;
        rl.pixels = scene->pixels;
        rl.max_depth = p.max_depth;
        rl.counters = scene->counters;
        rl.marks = p.use_graph ? nullptr : &scene->marks;
        rl.raster = raster ? &scene->raster : nullptr;
        rl.sraster = sraster ? &scene->sraster : nullptr;
        rl.side = ctx->side;
        GraphHolder gh;
        std::unique_ptr<PixelFeed> feed;
        if (cb && cb->pixels_after) feed = std::make_unique<PixelFeed>(ctx, scene, cb, batch_pixels);
        for (int64_t pass = 0; p.passes == -1 || pass < p.passes; ++pass) {
            if (stop && *stop) raise(NSP_E_STOPPED, "stopped");
            auto t0 = std::chrono::steady_clock::now();
            scene->marks.reset();
            pass_reset(ctx->work, scene->counters, scene->pass_dev.as<int>(), rng_pass(scene), s);
            if (p.use_graph && !gh.exec) {
                NSP_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                try {
                    render_batch(rl, s);
                } catch (...) {
                    hipGraph_t g2 = nullptr;
                    (void)hipStreamEndCapture(s, &g2);
                    if (g2) (void)hipGraphDestroy(g2);
                    throw;
                }
                NSP_HIP(hipStreamEndCapture(s, &gh.graph));
                NSP_HIP(hipGraphInstantiate(&gh.exec, gh.graph, nullptr, nullptr, 0));
            }
            int64_t done_pixels = 0;
            for (size_t b = 0; b < descs.size(); ++b) {
                if (stop && *stop) break;
                if (gh.exec)
                    NSP_HIP(hipGraphLaunch(gh.exec, s));
                else
                    render_batch(rl, s);
                if (feed) feed->batch_done(descs[b].pixel0, descs[b].paths / S, s);
                done_pixels += descs[b].paths / S;
            }
            if (feed) feed->deliver();
            unsigned long long tot[3] = {0, 0, 0};
            NSP_HIP(hipMemcpyAsync(tot, ctx->work.totals.p, sizeof tot, hipMemcpyDeviceToHost, s));
            NSP_HIP(hipStreamSynchronize(s));
            const unsigned long long closest = tot[0], shadow = tot[1], probes = tot[2];
            auto t1 = std::chrono::steady_clock::now();
            double secs = std::chrono::duration<double>(t1 - t0).count();
            // stage timing
            nsp_stage_stats st{};
            for (auto& [stage, e0, e1] : scene->marks.marks) {
                float ms = 0;
                NSP_HIP(hipEventElapsedTime(&ms, e0, e1));
                switch (stage) {
                case 0: st.ms_raygen += ms; break;
                case 1: st.ms_closest += ms; ++st.launches_closest; break;
                case 2: st.ms_shade += ms; break;
                case 3: st.ms_shadow += ms; ++st.launches_shadow; break;
                case 4: st.ms_accumulate += ms; break;
                case 5: st.ms_raster += ms; ++st.launches_raster; break;
                case 6: st.ms_walk += ms; ++st.launches_walk; break;
                }
            }
            st.ms_total = secs * 1e3;
            st.closest_rays = (int64_t)closest;
            st.primary_rays = raster ? scene->pixels * S : 0;
            st.shadow_rays = (int64_t)shadow;
            if (scene->counters) {
                unsigned long long tc[9];
                NSP_HIP(hipMemcpy(tc, ctx->work.trav_counters.p, sizeof tc, hipMemcpyDeviceToHost));
                st.node_visits = (int64_t)tc[0];
                st.leaf_visits = (int64_t)tc[1];
                st.simplex_tests = (int64_t)tc[2];
                st.retests = (int64_t)tc[3];
                st.shadow_node_visits = (int64_t)tc[4];
                st.shadow_leaf_visits = (int64_t)tc[5];
                st.shadow_simplex_tests = (int64_t)tc[6];
                st.shadow_retests = (int64_t)tc[7];
                st.shadow_rays_rasterised = (int64_t)tc[8];
            }
            scene->stage = st;
            bool stopped = stop && *stop;
            // BarPaintbrush statistics (paintbrush.h:177-232): pixels and samples
            // count as they are painted, while the pass count is the 1-based index
            // of the pass begun (m_pass_count starts at 1 and next_pass moves it
            // on, paintbrush.h:118, 200-219), so a stopped pass counts as begun;
            // pass_done follows only a pass whose batches all ran
            const bool whole = done_pixels == scene->pixels;
            scene->pass_count += 1;
            scene->pixel_count += done_pixels;
            scene->ray_count += (int64_t)(closest + shadow + probes);
            scene->sample_count += done_pixels * S;
            scene->last_pass_s = secs;
            if (whole && cb && cb->pass_done) cb->pass_done(cb->user, (int)scene->pass_count, secs);
            if (stopped) raise(NSP_E_STOPPED, "stopped");
        }
    });
    if (rc != NSP_OK && rc != NSP_E_STOPPED && cb && cb->error_message)
        cb->error_message(cb->user, (std::string("Painter error:\n") + ctx->error).c_str());
    return rc;
}

int nsp_statistics(const nsp_scene* scene, int64_t* pass_count, int64_t* pixel_count, int64_t* ray_count,
                   int64_t* sample_count, double* previous_pass_duration)
{
    if (!scene) return NSP_E_INVALID;
    if (pass_count) *pass_count = scene->pass_count;
    if (pixel_count) *pixel_count = scene->pixel_count;
    if (ray_count) *ray_count = scene->ray_count;
    if (sample_count) *sample_count = scene->sample_count;
    if (previous_pass_duration) *previous_pass_duration = scene->last_pass_s;
    return NSP_OK;
}

int nsp_image_get(const nsp_scene* scene, float* rgb_sum, int32_t* samples)
{
    if (!scene) return NSP_E_INVALID;
    return guarded(scene->ctx, [&] {
        const int64_t P = scene->pixels;
        if (rgb_sum) {
            std::vector<float> s((size_t)3 * P);
            NSP_HIP(hipMemcpy(s.data(), scene->pix_sum.p, sizeof(float) * 3 * P, hipMemcpyDeviceToHost));
            for (int64_t j = 0; j < P; ++j)
                for (int c = 0; c < 3; ++c) rgb_sum[j * 3 + c] = s[(size_t)c * P + j];
        }
        if (samples) NSP_HIP(hipMemcpy(samples, scene->pix_cnt.p, sizeof(int32_t) * P, hipMemcpyDeviceToHost));
    });
}

int nsp_image_set(nsp_scene* scene, const float* rgb_sum, const int32_t* samples)
{
    if (!scene || !rgb_sum || !samples) return NSP_E_INVALID;
    return guarded(scene->ctx, [&] {
        const int64_t P = scene->pixels;
        std::vector<float> s((size_t)3 * P);
        for (int64_t j = 0; j < P; ++j)
            for (int c = 0; c < 3; ++c) s[(size_t)c * P + j] = rgb_sum[j * 3 + c];
        NSP_HIP(hipMemcpy(scene->pix_sum.p, s.data(), sizeof(float) * 3 * P, hipMemcpyHostToDevice));
        NSP_HIP(hipMemcpy(scene->pix_cnt.p, samples, sizeof(int32_t) * P, hipMemcpyHostToDevice));
    });
}

int nsp_image_aov(const nsp_scene* scene, float* albedo_sum, float* normal_sum, float* direct_sum)
{
    if (!scene) return NSP_E_INVALID;
    return guarded(scene->ctx, [&] {
        const int64_t P = scene->pixels;
        const int n = scene->n;
        std::vector<float> a((size_t)pix_aov_channels(n) * P, 0.0f);
        if (scene->pix_aov.p)
            NSP_HIP(hipMemcpy(a.data(), scene->pix_aov.p, sizeof(float) * a.size(), hipMemcpyDeviceToHost));
        for (int64_t j = 0; j < P; ++j) {
            if (albedo_sum)
                for (int c = 0; c < 3; ++c) albedo_sum[j * 3 + c] = a[(size_t)c * P + j];
            if (normal_sum)
                for (int k = 0; k < n; ++k) normal_sum[j * n + k] = a[(size_t)(3 + k) * P + j];
            if (direct_sum)
                for (int c = 0; c < 3; ++c) direct_sum[j * 3 + c] = a[(size_t)(3 + n + c) * P + j];
        }
    });
}

int nsp_image_first_sample(const nsp_scene* scene, float* rgb_sum)
{
    if (!scene || !rgb_sum) return NSP_E_INVALID;
    return guarded(scene->ctx, [&] {
        const int64_t P = scene->pixels;
        std::vector<float> a((size_t)3 * P, 0.0f);
        if (scene->pix_aov.p)
            NSP_HIP(hipMemcpy(a.data(), scene->pix_aov.as<float>() + (size_t)aov_channels(scene->n) * P,
                              sizeof(float) * a.size(), hipMemcpyDeviceToHost));
        for (int64_t j = 0; j < P; ++j)
            for (int c = 0; c < 3; ++c) rgb_sum[j * 3 + c] = a[(size_t)c * P + j];
    });
}

int nsp_last_pass_stats(const nsp_scene* scene, nsp_stage_stats* out)
{
    if (!scene || !out) return NSP_E_INVALID;
    *out = scene->stage;
    return NSP_OK;
}

int nsp_set_counters(nsp_scene* scene, int enable)
{
    if (!scene) return NSP_E_INVALID;
    scene->counters = enable != 0;
    return NSP_OK;
}

}  // extern "C"

// A batch of scenes painted together on K lanes.  A lane is a stream with a
// side stream for the shadow stages and a workspace of its own (lane 0: the
// context's).  The scenes go to the lanes by longest processing time first
// over the cost estimate paths x log2(2 + simplices); each lane paints its
// scenes one after another, costliest first, with nsp_paint's wavefront
// batches (params->batch_paths, 0: the largest power of two of paths up to
// 2^27 within the lane's workspace budget), and the lanes run concurrently,
// so one scene's bounce tails and short stages overlap another scene's walks.
// Every scene keeps its own pass state (batch list, batch index, ray totals).
// With use_graph the whole batch is captured on the first call and its
// instantiated graph replayed by later calls with the same parameters, lanes,
// workspaces and scene constants; the pass indices are written before each
// replay.
struct BatchScene {
    DevBuf list, idx, totals;
    SceneK<float> kf;
    int64_t batch_pixels = 0, batches = 0;
    int lane = 0;
    bool sraster = false;  // the primary hits' shadow rays rasterised from the light
};

struct nsp_batch {
    nsp_context* ctx;
    int device = 0;
    int lanes = 2;
    std::vector<nsp_scene*> scenes;
    std::vector<BatchScene> per;
    // lanes 1..K-1: workspace, stream, side stream, join event
    std::vector<std::unique_ptr<RenderBuffers>> work;
    std::vector<hipStream_t> streams, sides;
    std::vector<hipEvent_t> joins;
    hipEvent_t fork = nullptr;
    // the cached graph and what it was captured for
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<int64_t> key;
    DevBuf pass_ptrs, pass_vals;  // each scene's pass_dev, and the pass indices written before a launch
    int64_t rays = 0;
    double last_s = 0;
    void drop_graph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        key.clear();
    }
};

namespace {
// the device addresses and sizes of a workspace's buffers: a captured graph
// holds them, so it is replayed only while they are unchanged
void fingerprint(const RenderBuffers& b, std::vector<int64_t>& out)
{
    auto add = [&](const DevBuf& d) {
        out.push_back((int64_t)(uintptr_t)d.p);
        out.push_back((int64_t)d.bytes);
    };
    for (const DevBuf& d : b.qa) add(d);
    for (const DevBuf& d : b.qb) add(d);
    for (const auto& q : b.sq)
        for (const DevBuf& d : q) add(d);
    for (const DevBuf* d : {&b.L, &b.Ls, &b.redo, &b.res, &b.keys, &b.redo_c, &b.res_c, &b.aov, &b.counters, &b.desc,
                            &b.batch_list, &b.batch_idx, &b.totals, &b.trav_counters, &b.sr_cell, &b.sr_rank,
                            &b.sr_rays, &b.sr_ray, &b.sr_occq})
        add(*d);
    out.push_back(b.cap);
    out.push_back(b.n);
    out.push_back(b.lights);
    out.push_back(b.depth_slots);
}

// what a captured segment of one scene holds: its buffers, raster work units
// and launch constants (FNV-1a over the constants' bytes)
void fingerprint(const nsp_scene& sc, const BatchScene& bs, std::vector<int64_t>& out)
{
    for (const DevBuf* d : {&bs.list, &bs.idx, &bs.totals, &sc.pass_dev, &sc.pix_sum, &sc.pix_cnt, &sc.raster.box,
                            &sc.raster.unit, &sc.raster.ustart, &sc.raster.planes, &sc.sraster.box,
                            &sc.sraster.planes, &sc.sraster.unit, &sc.sraster.count, &sc.sraster.start,
                            &sc.sraster.tiles}) {
        out.push_back((int64_t)(uintptr_t)d->p);
        out.push_back((int64_t)d->bytes);
    }
    const RasterPrep& r = sc.raster;
    for (int64_t v : {(int64_t)r.n, (int64_t)r.S, r.F, r.ppu, (int64_t)r.hmax, (int64_t)r.last, r.plane, r.units,
                      (int64_t)r.small, bs.batch_pixels, bs.batches, (int64_t)bs.lane, (int64_t)(uintptr_t)&sc.mesh->m,
                      (int64_t)bs.sraster, sc.sraster.cells, sc.sraster.units})
        out.push_back(v);
    uint64_t h = 1469598103934665603ull;
    const unsigned char* c = reinterpret_cast<const unsigned char*>(&bs.kf);
    for (size_t i = 0; i < sizeof(bs.kf); ++i) h = (h ^ c[i]) * 1099511628211ull;
    out.push_back((int64_t)h);
}

// writes pass_dev of every scene: ptrs[i] <- vals[i]
__global__ void k_set_passes(int* const* __restrict__ ptrs, const int* __restrict__ vals, int count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) *ptrs[i] = vals[i];
}
}  // namespace

extern "C" {

int nsp_batch_create(nsp_context* ctx, nsp_scene* const* scenes, int count, nsp_batch** out)
{
    if (!ctx || !out || !scenes || count < 1) return NSP_E_INVALID;
    *out = nullptr;
    return guarded(ctx, [&] {
        auto b = std::make_unique<nsp_batch>();
        b->ctx = ctx;
        b->device = ctx->device;
        for (int i = 0; i < count; ++i) {
            if (!scenes[i] || scenes[i]->ctx != ctx) raise(NSP_E_INVALID, "batch scenes must belong to the context");
            b->scenes.push_back(scenes[i]);
        }
        b->per.resize(count);
        NSP_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
        *out = b.release();
    });
}

void nsp_batch_destroy(nsp_batch* b)
{
    if (!b) return;
    (void)hipSetDevice(b->device);
    (void)hipDeviceSynchronize();
    b->drop_graph();
    for (hipStream_t st : b->streams) (void)hipStreamDestroy(st);
    for (hipStream_t st : b->sides) (void)hipStreamDestroy(st);
    for (hipEvent_t e : b->joins) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    delete b;
}

int nsp_batch_get_lanes(const nsp_batch* b) { return b ? b->lanes : NSP_E_INVALID; }

int nsp_batch_set_lanes(nsp_batch* b, int lanes)
{
    if (!b || lanes < 1 || lanes > 8) return NSP_E_INVALID;
    b->lanes = lanes;
    return NSP_OK;
}

int nsp_batch_paint(nsp_batch* b, const nsp_paint_params* params, const volatile int32_t* stop)
{
    if (!b || !params) return NSP_E_INVALID;
    nsp_context* ctx = b->ctx;
    return guarded(ctx, [&] {
        const nsp_paint_params& p = *params;
        if (p.precision != NSP_PRECISION_FP32) raise(NSP_E_UNSUPPORTED, "the batched launch traces in FP32");
        if (p.max_depth < 0 || p.max_depth > 250) raise(NSP_E_INVALID, "max_depth must be in [0, 250]");
        if (p.aov) raise(NSP_E_UNSUPPORTED, "AOVs are not produced by the batched launch");
        if (p.primary != NSP_PRIMARY_AUTO && p.primary != NSP_PRIMARY_WALK) raise(NSP_E_INVALID, "unknown primary mode");
        launch_gate(ctx);
        hipStream_t s = ctx->stream;
        const int ns = (int)b->scenes.size();
        const int lanes = std::min(b->lanes, ns);
        const bool raster_mode = p.primary == NSP_PRIMARY_AUTO;
        // each lane's workspace budget: the single-scene budget for up to two
        // lanes, twice that shared among more
        const int64_t budget = lanes <= 2 ? DEFAULT_WORKSPACE : 2 * DEFAULT_WORKSPACE / lanes;
        while ((int)b->streams.size() < lanes - 1) {
            hipStream_t st, sd;
            NSP_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            b->streams.push_back(st);
            NSP_HIP(hipStreamCreateWithFlags(&sd, hipStreamNonBlocking));
            b->sides.push_back(sd);
            hipEvent_t e;
            NSP_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            b->joins.push_back(e);
            b->work.push_back(std::make_unique<RenderBuffers>());
        }
        auto lane_work = [&](int l) -> RenderBuffers& { return l == 0 ? ctx->work : *b->work[l - 1]; };
        auto lane_stream = [&](int l) { return l == 0 ? s : b->streams[l - 1]; };
        auto lane_side = [&](int l) { return l == 0 ? ctx->side : b->sides[l - 1]; };
        // per scene: launch constants, cost, batch rule
        std::vector<double> cost(ns);
        for (int i = 0; i < ns; ++i) {
            nsp_scene* sc = b->scenes[i];
            BatchScene& bs = b->per[i];
            if (!sc->shapes.empty() && p.smooth_normal && sc->mesh->m.has_normals)
                raise(NSP_E_UNSUPPORTED, "smooth-normal renders of meshes with vertex normals do not support shapes");
            fill_scene<float>(*sc, p, 0, bs.kf);
            if (!sc->pass_dev.p) sc->pass_dev.alloc(sizeof(int));
            bs.kf.pass_dev = sc->pass_dev.as<int>();
            const int S = bs.kf.S;
            cost[i] = (double)(sc->pixels * S) * std::log2(2.0 + (double)sc->mesh->m.F);
            const bool raster = raster_mode && sc->proj.kind == NSP_PROJECTOR_PARALLEL;
            // as in nsp_paint; shadow_raster_prepare may still turn it off below
            bs.sraster = raster_mode && sc->lights.size() == 1 && sc->lights[0].kind == NSP_LIGHT_POINT &&
                         sc->shapes.empty() && sc->mesh->m.layout == NSP_LAYOUT_SIMPLEX_AOS &&
                         !(p.smooth_normal && sc->mesh->m.has_normals);
            int64_t want = p.batch_paths;
            if (want <= 0) {
                const int64_t per = render_bytes_per_path(sc->n, NSP_PRECISION_FP32, (int)sc->lights.size(), false,
                                                          raster, bs.sraster);
                want = (int64_t)1 << 27;
                while (want > ((int64_t)1 << 20) && want * per > budget) want >>= 1;
            }
            bs.batch_pixels = std::min<int64_t>(std::max<int64_t>(1, want / S), sc->pixels);
        }
        // lanes by longest processing time first
        std::vector<int> order(ns);
        for (int i = 0; i < ns; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return cost[a] > cost[c]; });
        std::vector<double> load(lanes, 0.0);
        std::vector<std::vector<int>> lane_scenes(lanes);
        for (int i : order) {
            const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            b->per[i].lane = l;
            lane_scenes[l].push_back(i);
            load[l] += cost[i];
        }
        // workspaces (they only grow), raster set-up, batch lists
        for (int l = 0; l < lanes; ++l) {
            int64_t cap = 1, srcap = 0;
            int nmax = 3, lmax = 1;
            for (int i : lane_scenes[l]) {
                const int64_t c = b->per[i].batch_pixels * b->per[i].kf.S;
                cap = std::max(cap, c);
                if (b->per[i].sraster) srcap = std::max(srcap, c);
                nmax = std::max(nmax, b->scenes[i]->n);
                lmax = std::max(lmax, (int)b->scenes[i]->lights.size());
            }
            alloc_render_buffers(lane_work(l), nmax, NSP_PRECISION_FP32, cap, p.max_depth, lmax);
            if (raster_mode) lane_work(l).keys.alloc(sizeof(unsigned long long) * cap);
            if (srcap > 0) shadow_raster_buffers(lane_work(l), nmax, srcap);
        }
        for (int i = 0; i < ns; ++i) {
            nsp_scene* sc = b->scenes[i];
            BatchScene& bs = b->per[i];
            const int S = bs.kf.S;
            std::vector<BatchDesc> descs;
            for (int64_t px = 0; px < sc->pixels; px += bs.batch_pixels)
                descs.push_back(BatchDesc{px, std::min(bs.batch_pixels, sc->pixels - px) * S});
            bs.batches = (int64_t)descs.size();
            bs.list.alloc(sizeof(BatchDesc) * descs.size());
            bs.idx.alloc(sizeof(long long));
            bs.totals.alloc(sizeof(unsigned long long) * 3);
            NSP_HIP(hipMemcpyAsync(bs.list.p, descs.data(), sizeof(BatchDesc) * descs.size(), hipMemcpyHostToDevice,
                                   s));
            if (raster_mode && (sc->proj.kind == NSP_PROJECTOR_PARALLEL || bs.sraster)) {
                SceneK<double> kd;
                fill_scene<double>(*sc, p, 0, kd);
                if (sc->proj.kind == NSP_PROJECTOR_PARALLEL) raster_prepare(sc->raster, sc->mesh->m, kd, S, s);
                if (bs.sraster) bs.sraster = shadow_raster_prepare(sc->sraster, sc->mesh->m, kd.light_pos[0], s);
            }
        }
        // the pass index of every scene, written before each launch
        std::vector<int> vals(ns);
        std::vector<int*> ptrs(ns);
        for (int i = 0; i < ns; ++i) {
            vals[i] = rng_pass(b->scenes[i]);
            ptrs[i] = b->scenes[i]->pass_dev.as<int>();
        }
        b->pass_ptrs.alloc(sizeof(int*) * ns);
        b->pass_vals.alloc(sizeof(int) * ns);
        NSP_HIP(hipMemcpyAsync(b->pass_ptrs.p, ptrs.data(), sizeof(int*) * ns, hipMemcpyHostToDevice, s));
        NSP_HIP(hipMemcpyAsync(b->pass_vals.p, vals.data(), sizeof(int) * ns, hipMemcpyHostToDevice, s));
        NSP_HIP(hipStreamSynchronize(s));
        if (stop && *stop) raise(NSP_E_STOPPED, "stopped");
        auto launch_for = [&](int i) {
            nsp_scene* sc = b->scenes[i];
            BatchScene& bs = b->per[i];
            RenderLaunch rl;
            rl.mesh = &sc->mesh->m;
            rl.scene = &bs.kf;
            rl.buf = &lane_work(bs.lane);
            rl.pix_sum = sc->pix_sum.as<float>();
            rl.pix_cnt = sc->pix_cnt.as<int32_t>();
            rl.pix_aov = nullptr;
            rl.pixels = sc->pixels;
            rl.max_depth = p.max_depth;
            rl.counters = false;
            rl.marks = nullptr;
            rl.raster = raster_mode && sc->proj.kind == NSP_PROJECTOR_PARALLEL ? &sc->raster : nullptr;
            rl.sraster = bs.sraster ? &sc->sraster : nullptr;
            rl.batch_list = bs.list.as<BatchDesc>();
            rl.batch_idx = bs.idx.as<long long>();
            rl.totals = bs.totals.as<unsigned long long>();
            rl.side = lane_side(bs.lane);
            return rl;
        };
        // the lanes' launch sequences, submitted round robin one wavefront
        // batch at a time so that no lane waits for another's submission
        auto enqueue = [&]() {
            NSP_HIP(hipEventRecord(b->fork, s));
            for (int l = 1; l < lanes; ++l) NSP_HIP(hipStreamWaitEvent(lane_stream(l), b->fork, 0));
            std::vector<size_t> next(lanes, 0);
            std::vector<int64_t> left(lanes, 0);
            std::vector<RenderLaunch> cur(lanes);
            for (bool more = true; more;) {
                more = false;
                for (int l = 0; l < lanes; ++l) {
                    if (left[l] == 0) {
                        if (next[l] == lane_scenes[l].size()) continue;
                        const int i = lane_scenes[l][next[l]++];
                        pass_reset_state(b->per[i].totals.as<unsigned long long>(), b->per[i].idx.as<long long>(),
                                         nullptr, 0, lane_stream(l));
                        cur[l] = launch_for(i);
                        left[l] = b->per[i].batches;
                    }
                    render_batch(cur[l], lane_stream(l));
                    --left[l];
                    more = true;
                }
            }
            for (int l = 1; l < lanes; ++l) {
                NSP_HIP(hipEventRecord(b->joins[l - 1], lane_stream(l)));
                NSP_HIP(hipStreamWaitEvent(s, b->joins[l - 1], 0));
            }
        };
        if (p.use_graph) {
            std::vector<int64_t> key = {p.samples_per_pixel, p.max_depth, p.smooth_normal, (int64_t)p.seed,
                                        p.batch_paths, p.primary, lanes};
            for (int l = 0; l < lanes; ++l) fingerprint(lane_work(l), key);
            for (int i = 0; i < ns; ++i) fingerprint(*b->scenes[i], b->per[i], key);
            if (!b->exec || key != b->key) {
                b->drop_graph();
                NSP_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                try {
                    enqueue();
                } catch (...) {
                    hipGraph_t g2 = nullptr;
                    (void)hipStreamEndCapture(s, &g2);
                    if (g2) (void)hipGraphDestroy(g2);
                    throw;
                }
                NSP_HIP(hipStreamEndCapture(s, &b->graph));
                NSP_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
                b->key = key;
            }
        }
        auto t0 = std::chrono::steady_clock::now();
        hipLaunchKernelGGL(k_set_passes, dim3((ns + 255) / 256), dim3(256), 0, s, b->pass_ptrs.as<int*>(),
                           b->pass_vals.as<int>(), ns);
        NSP_HIP(hipGetLastError());
        if (p.use_graph)
            NSP_HIP(hipGraphLaunch(b->exec, s));
        else
            enqueue();
        NSP_HIP(hipStreamSynchronize(s));
        b->last_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        b->rays = 0;
        std::vector<unsigned long long> tot((size_t)3 * ns);
        for (int i = 0; i < ns; ++i)
            NSP_HIP(hipMemcpyAsync(tot.data() + 3 * i, b->per[i].totals.p, 3 * sizeof(unsigned long long),
                                   hipMemcpyDeviceToHost, s));
        NSP_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < ns; ++i) {
            nsp_scene* sc = b->scenes[i];
            const int64_t r = (int64_t)(tot[3 * i] + tot[3 * i + 1] + tot[3 * i + 2]);
            b->rays += r;
            sc->pass_count += 1;
            sc->pixel_count += sc->pixels;
            sc->ray_count += r;
            sc->sample_count += sc->pixels * b->per[i].kf.S;
            sc->last_pass_s = b->last_s;
        }
    });
}

int nsp_batch_statistics(const nsp_batch* b, int64_t* ray_count, double* last_pass_duration)
{
    if (!b) return NSP_E_INVALID;
    if (ray_count) *ray_count = b->rays;
    if (last_pass_duration) *last_pass_duration = b->last_s;
    return NSP_OK;
}

int nsp_autotune(nsp_scene* scene, nsp_paint_params* params)
{
    if (!scene || !params) return NSP_E_INVALID;
    nsp_context* ctx = scene->ctx;
    // save the scene state, time candidate queue sizes on whole passes, restore
    const int64_t P = scene->pixels;
    std::vector<float> sum((size_t)3 * P);
    std::vector<int32_t> cnt(P);
    int rc = nsp_image_get(scene, sum.data(), cnt.data());
    if (rc != NSP_OK) return rc;
    const int64_t passes = scene->pass_count, pixels = scene->pixel_count, rays = scene->ray_count,
                  samples = scene->sample_count;
    const double last = scene->last_pass_s;
    nsp_paint_params p = *params;
    p.passes = 1;
    // the timed passes write no AOVs: the AOV sums of the scene stay as they are
    // (the queue size does not depend on them beyond their buffer, which the
    // chosen size allocates on the caller's first paint with AOVs)
    p.aov = 0;
    int64_t best = 0;
    double best_s = 1e300;
    const bool raster = p.precision == NSP_PRECISION_FP32 && p.primary == NSP_PRIMARY_AUTO &&
                        scene->proj.kind == NSP_PROJECTOR_PARALLEL;
    const int64_t per = render_bytes_per_path(scene->n, p.precision, (int)scene->lights.size(), false, raster);
    int64_t S = 0;
    rc = guarded(ctx, [&] {
        S = 1;
        const int k = one_dimension_size(scene->n - 1, p.samples_per_pixel);
        for (int i = 0; i < scene->n - 1; ++i) S *= k;
    });
    if (rc != NSP_OK) return rc;
    const int64_t all_paths = scene->pixels * S;
    // candidates: 2^20 .. the first power of two that holds the whole screen in
    // one batch (larger ones give the same batches), within the workspace budget
    std::vector<int64_t> cand;
    for (int lg = 20; lg <= (p.precision ? 26 : 27); ++lg) {
        const int64_t paths = (int64_t)1 << lg;
        if (lg > 20 && paths * per > DEFAULT_WORKSPACE) break;
        cand.push_back(paths);
        if (paths >= all_paths) break;
    }
    // untimed warm-up at the largest candidate: the workspace (which only grows)
    // and the raster set-up are then in place before any candidate is timed
    p.batch_paths = cand.back();
    rc = nsp_paint(scene, &p, nullptr, nullptr);
    for (size_t c = 0; rc == NSP_OK && c < cand.size(); ++c) {
        p.batch_paths = cand[c];
        auto t0 = std::chrono::steady_clock::now();
        int r = nsp_paint(scene, &p, nullptr, nullptr);
        if (r != NSP_OK) {
            // a failed candidate ends the search; the best size so far stands
            if (!best) rc = r;
            break;
        }
        double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (secs < best_s) {
            best_s = secs;
            best = p.batch_paths;
        }
    }
    int rc2 = nsp_image_set(scene, sum.data(), cnt.data());
    scene->pass_count = passes;
    scene->pixel_count = pixels;
    scene->ray_count = rays;
    scene->sample_count = samples;
    scene->last_pass_s = last;
    if (rc != NSP_OK) return rc;
    if (rc2 != NSP_OK) return rc2;
    params->batch_paths = best;
    return NSP_OK;
}

// Batched launch: the lane count (1, 2, 3, 4), each timed on one whole batched
// pass after an untimed pass (which captures the graph with use_graph);
// candidates whose workspaces cannot be allocated are skipped; every scene's
// image and statistics are restored.
int nsp_batch_autotune(nsp_batch* b, nsp_paint_params* params)
{
    if (!b || !params) return NSP_E_INVALID;
    const int ns = (int)b->scenes.size();
    struct Saved {
        std::vector<float> sum;
        std::vector<int32_t> cnt;
        int64_t passes, pixels, rays, samples;
        double last;
    };
    std::vector<Saved> saved(ns);
    for (int i = 0; i < ns; ++i) {
        nsp_scene* sc = b->scenes[i];
        Saved& sv = saved[i];
        sv.sum.resize((size_t)3 * sc->pixels);
        sv.cnt.resize(sc->pixels);
        int rc = nsp_image_get(sc, sv.sum.data(), sv.cnt.data());
        if (rc != NSP_OK) return rc;
        sv.passes = sc->pass_count;
        sv.pixels = sc->pixel_count;
        sv.rays = sc->ray_count;
        sv.samples = sc->sample_count;
        sv.last = sc->last_pass_s;
    }
    nsp_paint_params p = *params;
    p.passes = 1;
    int rc = NSP_OK;
    const int lanes0 = b->lanes;
    int best_lanes = 0;
    double best_s = 1e300;
    for (int lanes : {1, 2, 3, 4}) {
        if (rc != NSP_OK) break;
        if (lanes > ns) continue;
        b->lanes = lanes;
        rc = nsp_batch_paint(b, &p, nullptr);
        double t = 1e300;
        if (rc == NSP_OK) {
            auto t0 = std::chrono::steady_clock::now();
            rc = nsp_batch_paint(b, &p, nullptr);
            t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
        if (rc == NSP_E_NOMEM) {
            (void)hipGetLastError();
            rc = NSP_OK;
            continue;
        }
        if (rc == NSP_OK && t < best_s) {
            best_s = t;
            best_lanes = lanes;
        }
    }
    if (rc == NSP_OK && !best_lanes) {
        rc = NSP_E_NOMEM;
        b->ctx->error = "no batch configuration fits the device memory";
    }
    int rc2 = NSP_OK;
    for (int i = 0; i < ns; ++i) {
        nsp_scene* sc = b->scenes[i];
        const Saved& sv = saved[i];
        int r = nsp_image_set(sc, sv.sum.data(), sv.cnt.data());
        if (r != NSP_OK) rc2 = r;
        sc->pass_count = sv.passes;
        sc->pixel_count = sv.pixels;
        sc->ray_count = sv.rays;
        sc->sample_count = sv.samples;
        sc->last_pass_s = sv.last;
    }
    if (rc != NSP_OK || rc2 != NSP_OK) {
        b->lanes = lanes0;
        return rc != NSP_OK ? rc : rc2;
    }
    b->lanes = best_lanes;
    return NSP_OK;
}

}  // extern "C"
