// nsp_render.hpp -- host/device interface of the trace and render kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <tuple>
#include <vector>

#include "nsp_internal.hpp"
#include "nsp_trace.hpp"

namespace nsp {

// Light sources per scene.  With one light a path has at most one shadow ray
// per bounce and the shadow stage adds to its radiance without atomics; with
// more, the additions are atomic.
constexpr int MAX_LIGHTS = 8;

// A flat-shaded shape beside the mesh, FP64 (VisibleParallelotope /
// VisibleHyperplaneParallelotope, src/painter/visible_shapes.h:31-117).
struct ShapeRec {
    int kind;                // NSP_SHAPE_*
    int light;               // a light source (emits light_rgb to diffuse bounces)
    float rgb[3], light_rgb[3], diffuse;
    double org[NMAX];
    double n[NMAX];          // hyperplane parallelotope: unit normal
    double pn[NMAX][NMAX];   // parallelotope: outward unit plane normals; hyperplane: barycentric planes
    double d1[NMAX], d2[NMAX];  // parallelotope: offsets of the two planes; hyperplane: d1 of the barycentric planes
};

// Scene constants passed by value to the render kernels (the data of
// PaintObjects: src/painter/objects.h:212-223, with the derived constants of
// paint_threads, src/painter/painter.cpp:560-575).
template <typename T>
struct SceneK {
    int proj;                // NSP_PROJECTOR_*
    T proj_r2;               // spherical: squared radius of the sphere the screen is projected onto
    T cam_org[NMAX];
    T cam_dir[NMAX];         // unit; perspective: times the distance of the screen (in pixels)
    T axes[NMAX - 1][NMAX];  // unit screen axes; parallel: times units_per_pixel
    int screen[NMAX - 1];
    int64_t pixels;
    int k;        // samples per screen dimension
    int S;        // samples per pixel = k^(n-1)
    T k_inv;      // 1/k in T
    int light_count;
    int light_kind[MAX_LIGHTS];
    T light_pos[MAX_LIGHTS][NMAX];
    float light_rgb[MAX_LIGHTS][3];
    T light_coef[MAX_LIGHTS];
    float bg_rgb[3];
    float bg_light_rgb[3];
    int bg_is_light;
    float obj_rgb[3];
    float obj_light_rgb[3];
    int obj_is_light;
    float diffuse;
    T ray_offset;
    T cos_coef;   // cosine_sphere_coefficient(n)
    int max_depth;
    int smooth;
    int has_normals;
    uint64_t seed;
    const ShapeRec* shapes;  // the objects after the mesh, in order (device memory)
    int shape_count;
    int pass;             // pass index, used when pass_dev is null
    const int* pass_dev;  // pass index in device memory (written by the pass reset), so that
                          // one captured launch sequence serves every pass
};

// The pass index of the sample streams.
template <typename S>
__device__ __forceinline__ int pass_index(const S& sc)
{
    return sc.pass_dev ? *sc.pass_dev : sc.pass;
}

struct BatchDesc {
    int64_t pixel0;
    int64_t paths;
};

// Ray-state queue (SoA, component-major: org[c * cap + i]).
template <typename T>
struct RayQueue {
    T* org;
    T* dir;
    int32_t* path;
    float* beta;   // 3 * cap
    float* level;  // colour level (painter.cpp:370-373)
    int32_t* hit;
    T* hit_t;
};

// Per-path AOV channels: first-hit albedo (3), first-hit normal (N) and the
// radiance the first hit receives directly (3: background or emitter seen by
// the primary ray plus the unoccluded first-hit light contributions).
__host__ __device__ constexpr int aov_channels(int n) { return 6 + n; }
// Per-pixel AOV sums: the per-path channels, then the radiance of each pass's
// first sample of the pixel (3: one sample per pixel and pass).
__host__ __device__ constexpr int pix_aov_channels(int n) { return aov_channels(n) + 3; }

template <typename T>
struct ShadowQueue {
    T* org;
    T* dir;
    T* tmax;
    int32_t* path;
    float* contrib;  // 3 * cap
    // with shapes: the first shape (in object order) that blocks the light, or -1
    int32_t* shape;
    int shape_count;
    unsigned long long* extra;  // ray counts of the shape tests (the mesh test is the shadow ray itself)
    int multi;                  // several shadow rays per path (more than one light): atomic additions
};

struct MeshShadeArgs {
    const float* normal32;
    const double* rec64;
    const double* vn64;
    const int32_t* fnrm;
    const int32_t* ntype;
};

// Queue segments.  Device-scope atomics on one address serialise (about 12 ns
// each on MI355X, against under 2 ns when eight addresses share the load), so
// a stage that appends to a queue with one atomic per wave would be bound by
// that counter.  Every ray and shadow queue is therefore split into QSEG
// segments, one per XCD: segment k holds its rays at [k * segcap, k * segcap +
// count_k) and has a counter of its own.  The shade stage's block b reads input
// segment b % QSEG and appends to output segment b % QSEG, so a segment never
// receives more rays than its input segment held; the persistent walks take
// segment k as the slice of the blocks on XCD k.
constexpr int QSEG = 8;
constexpr int CNT_LINE = 16;  // u64 per 128-byte line: one line per atomically updated counter

// Per-depth counters, CNT_STRIDE u64 per depth.  Line 0 holds the scalars:
constexpr int CNT_RAYS = 0;      // rays of the depth's input queue (depth 0: the batch's paths)
constexpr int CNT_PROBE = 2;     // smooth-normal probe rays
constexpr int CNT_RES_C = 3;     // closest-hit walks with pending FP64 decisions
constexpr int CNT_RES_S = 4;     // shadow walks with pending FP64 decisions
constexpr int CNT_REDO_C = 5;    // closest-hit re-walks
constexpr int CNT_REDO_S = 6;    // shadow re-walks
constexpr int CNT_EXTRA = 7;     // ray counts of shadow rays' shape tests
constexpr int CNT_SEGCAP = 8;    // segment capacity of the depth's input ray queue
constexpr int CNT_SSEGCAP = 9;   // segment capacity of the depth's shadow queue
// then one line per segment counter:
__host__ __device__ constexpr int cnt_rayseg(int k) { return CNT_LINE * (1 + k); }              // input rays
__host__ __device__ constexpr int cnt_shadowseg(int k) { return CNT_LINE * (1 + QSEG + k); }    // shadow rays
__host__ __device__ constexpr int cnt_cwork(int k) { return CNT_LINE * (1 + 2 * QSEG + k); }    // closest-hit walks taken
__host__ __device__ constexpr int cnt_swork(int k) { return CNT_LINE * (1 + 3 * QSEG + k); }    // shadow walks taken
constexpr int CNT_STRIDE = CNT_LINE * (1 + 4 * QSEG);

struct RenderBuffers {
    int n = 0;
    int precision = 0;
    int64_t cap = 0;        // paths per batch, a multiple of QSEG (queue arrays are component-major with stride cap)
    int lights = 1;         // shadow rays per path and bounce the shadow queues hold
    int depth_slots = 0;
    DevBuf qa[8], qb[8];
    DevBuf sq[2][6];        // shadow-ray queues of even and odd depths (ray, tmax, path, contribution, blocking shape)
    DevBuf L;               // 3 * cap floats: radiance the paths collect at their hits (background, emitters)
    DevBuf Ls;              // 3 * cap floats: radiance of the paths' unoccluded shadow rays
    // ordering of the shadow stages on a second stream (see render_batch)
    hipEvent_t fork_ev = nullptr, join_ev = nullptr, done_ev[2] = {nullptr, nullptr};
    RenderBuffers() = default;
    RenderBuffers(const RenderBuffers&) = delete;
    RenderBuffers& operator=(const RenderBuffers&) = delete;
    ~RenderBuffers();
    DevBuf redo;            // int32 shadow-ray indices to walk again with the FP64 re-test inline
    DevBuf res;             // int32 (shadow ray, PENDING candidates) for the FP64 decisions
    DevBuf keys;            // u64 per path: (t, simplex) of the rasterised primary hits
    DevBuf redo_c, res_c;   // the same two lists for closest-hit walks of bounce rays
    DevBuf aov;             // aov_channels(n) * cap floats per path
    DevBuf counters;        // (max_depth + 2) * CNT_STRIDE u64
    DevBuf desc;            // BatchDesc of the batch in flight
    DevBuf batch_list;      // BatchDesc per batch of the pass
    DevBuf batch_idx;       // long long: index of the next batch
    DevBuf totals;          // 3 u64: closest, shadow and probe rays of the pass
    DevBuf trav_counters;   // 9 u64: traversal counters (closest stage, shadow stage), rasterised shadow rays
    // binned shadow rays of the light-space rasteriser (nsp_raster.hpp)
    DevBuf sr_cell;         // int32 per shadow ray: its cell
    DevBuf sr_rank;         // u32 per shadow ray: its place among its cell's rays
    DevBuf sr_rays;         // per shadow ray in cell order: origin, direction, limit (negative once occluded)
    DevBuf sr_ray;          // int32 per binned ray: its shadow-queue index
    DevBuf sr_occq;         // u8 per shadow ray in queue order: occluded
};

// Adds an unoccluded shadow ray's contribution to its path (and, for a
// first-hit shadow ray while AOVs are collected, to the direct-light AOV).
// With one light, each path receives at most one contribution per shadow
// launch and the additions need no atomics; with several they are atomic.
template <int N, typename T>
__device__ __forceinline__ void shadow_unoccluded(const ShadowQueue<T>& sq, int64_t scap, int64_t i,
                                                  float* __restrict__ L, int64_t cap, float* __restrict__ direct)
{
    if (sq.shape_count > 0) {
        // the mesh (the first object) let the ray pass: the shapes were tested
        // next, up to the first that blocks the light (shade recorded it)
        const int32_t j = sq.shape[i];
        atomicAdd(sq.extra, (unsigned long long)(j >= 0 ? j + 1 : sq.shape_count));
        if (j >= 0) return;
    }
    const int32_t raw = sq.path[i];
    const int32_t path = raw < 0 ? -raw - 1 : raw;
    const float c0 = sq.contrib[i], c1 = sq.contrib[scap + i], c2 = sq.contrib[2 * scap + i];
    if (sq.multi) {
        atomicAdd(&L[path], c0);
        atomicAdd(&L[cap + path], c1);
        atomicAdd(&L[2 * cap + path], c2);
        if (raw < 0 && direct) {
            atomicAdd(&direct[path], c0);
            atomicAdd(&direct[cap + path], c1);
            atomicAdd(&direct[2 * cap + path], c2);
        }
        return;
    }
    L[path] += c0;
    L[cap + path] += c1;
    L[2 * cap + path] += c2;
    if (raw < 0 && direct) {
        direct[path] += c0;
        direct[cap + path] += c1;
        direct[2 * cap + path] += c2;
    }
}

TreeArgs make_tree_args(const MeshDev& m);

// First hit of a batch of rays (device pointers).
void launch_first_hit(const MeshDev& m, const float* org, const float* dir, int64_t count, int32_t* idx, double* t,
                      uint8_t* near, unsigned long long* counters, hipStream_t s);

// HIP events around each stage launch (stage ids: 0 raygen, 1 closest,
// 2 shade, 3 shadow, 4 accumulate, 5 rasterised primary visibility, 6 merged
// closest-hit and shadow walks).
struct StageMarks {
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    std::vector<std::tuple<int, hipEvent_t, hipEvent_t>> marks;
    hipEvent_t get();
    void reset()
    {
        used = 0;
        marks.clear();
    }
    ~StageMarks();
};

// One wavefront batch of the render (all kernels enqueued on s).
struct RasterPrep;

struct RenderLaunch {
    const MeshDev* mesh;
    const void* scene;  // SceneK<T>*
    RenderBuffers* buf;
    float* pix_sum;     // 3 * pixels (component-major: sum[c * pixels + p])
    int32_t* pix_cnt;
    float* pix_aov;     // (3 + n) * pixels sums of first-hit albedo and normal, or null
    int64_t pixels;
    int max_depth;
    bool counters;
    StageMarks* marks;  // may be null
    const RasterPrep* raster = nullptr;  // FP32 only: rasterised depth-0 hits (nsp_raster.hpp), or null to walk
    // a scene's own pass state when several scenes share one workspace (null: the workspace's)
    const BatchDesc* batch_list = nullptr;
    long long* batch_idx = nullptr;
    unsigned long long* totals = nullptr;
    // a second stream for the shadow stages (null: everything on the launch stream)
    hipStream_t side = nullptr;
    // FP32 walks over simplex-ordered records: the shadow rays rasterised from
    // the light (nsp_raster.hpp), or null to walk them
    const struct ShadowRasterPrep* sraster = nullptr;
};
// One wavefront batch: raygen, then per depth closest hits, shade and the
// shadow stage, then accumulate.  With rl.side, the shadow stage of depth d
// runs on the side stream, overlapping the closest hits and shading of depth
// d+1: it reads the shadow queue of depth d's parity and writes only Ls, and
// shade(d+2), which refills that queue, waits for it.  The side stream joins
// the launch stream before accumulate, so a caller sees one stream's work
// (and a graph captured on it holds both).
void render_batch(const RenderLaunch& rl, hipStream_t s);
// Clears the pass totals, the batch index and (when counting) the traversal counters,
// and writes the pass index to pass_dev (when not null).
void pass_reset(RenderBuffers& b, bool counters, int* pass_dev, int pass, hipStream_t s);
// Clears a scene's own pass totals (3 u64) and batch index.
void pass_reset_state(unsigned long long* totals, long long* batch_idx, int* pass_dev, int pass, hipStream_t s);
// Device bytes per path of one wavefront batch's workspace.
int64_t render_bytes_per_path(int n, int precision, int lights, bool aov, bool raster, bool shadow_raster = false);
void alloc_render_buffers(RenderBuffers& b, int n, int precision, int64_t cap, int max_depth, int lights = 1);
// FP64 records of shapes (org and edge vectors as given) into out (device), on stream s.
void build_shape_records(int n, const int* kind, const double* org, const double* edges, int count, ShapeRec* out,
                         hipStream_t s);

}  // namespace nsp
