// nsp_internal.hpp -- host-side internal structures of libnsp (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "nsp.h"

namespace nsp {

// Error raised inside the library; converted to a status code at the C ABI.
struct Error {
    int code;
    std::string msg;
};

[[noreturn]] void raise(int code, const std::string& msg);
void check_hip(hipError_t e, const char* what);
#define NSP_HIP(x) ::nsp::check_hip((x), #x)

// Device buffer owned by RAII.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes)
    {
        o.p = nullptr;
        o.bytes = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t b)
    {
        if (b <= bytes && p) return;
        release();
        if (b == 0) b = 16;
        NSP_HIP(hipMalloc(&p, b));
        bytes = b;
    }
    template <typename T>
    T* as() const
    {
        return static_cast<T*>(p);
    }
};

// Pinned host memory that only grows (staging of device-to-host copies that
// overlap later kernels).
struct HostBuf {
    void* p = nullptr;
    size_t bytes = 0;
    HostBuf() = default;
    HostBuf(const HostBuf&) = delete;
    HostBuf& operator=(const HostBuf&) = delete;
    ~HostBuf() { release(); }
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t b)
    {
        if (b <= bytes && p) return;
        release();
        NSP_HIP(hipHostMalloc(&p, b, hipHostMallocDefault));
        bytes = b;
    }
    template <typename T>
    T* as() const
    {
        return static_cast<T*>(p);
    }
};

// Per-simplex FP64 record layout (AoS, original simplex order):
//   normal[N] | vertex0[N] | plane_n[(N-1)*N] | plane_d[N-1]
__host__ __device__ constexpr int rec64_stride(int n) { return n * n + 2 * n - 1; }
// Per-simplex FP64 overlap-test record: smin[N] | smax[N] | cons[N*(N+1)] | ceq[N+1]
__host__ __device__ constexpr int build64_stride(int n) { return 2 * n + n * (n + 1) + n + 1; }
// FP32 leaf-ordered SoA components: normal[N], d, then (plane_n[N], plane_d) * (N-1)
__host__ __device__ constexpr int soa_components(int n) { return n * (n + 1); }
// Per-simplex FP32 record (AoS, simplex order), the same components padded to
// a multiple of 4 floats so that a record is read with 16-byte loads.
__host__ __device__ constexpr int rec32_stride(int n) { return (n * (n + 1) + 3) / 4 * 4; }

struct TreeDev {
    int n = 0;
    int max_depth = 0;      // reference depth (root depth 1)
    int levels = 0;         // number of levels actually built
    double root_lo[NSP_MAX_DIM] = {0};
    double root_size[NSP_MAX_DIM] = {0};
    double dist_from_facet = 0;
    int64_t node_count = 0;
    int64_t entry_count = 0;
    int64_t leaf_count = 0;
    DevBuf nodes;       // int2 per node: internal {first child, -stride}; leaf {entry begin, count}
    DevBuf node_level;  // int8 level per node (for dumps / statistics)
    DevBuf node_q;      // int32 q[N] per node (for dumps)
    DevBuf entries;     // int32 simplex index per leaf entry
    DevBuf soa;         // float, soa_components(n) arrays of entry_pad
    int64_t entry_pad = 0;
};

// A scene's per-pixel sums in device memory (nsp_api.cpp keeps the scene):
// sum [3][pixels], cnt [pixels], aov [pix_aov_channels(n)][pixels].
struct ScenePixels {
    nsp_context* ctx = nullptr;
    int n = 0;
    int64_t pixels = 0;
    int screen[NSP_MAX_DIM - 1] = {};
    const float* sum = nullptr;
    const int32_t* cnt = nullptr;
    const float* aov = nullptr;
};
bool scene_pixel_buffers(const nsp_scene* s, ScenePixels* out);

struct MeshDev {
    int n = 0;
    int64_t V = 0, F = 0, VN = 0;
    bool has_normals = false;
    double mn[NSP_MAX_DIM], mx[NSP_MAX_DIM];
    DevBuf vtx64;     // V*n double
    DevBuf fac;       // F*n int32
    DevBuf vn64;      // VN*n double (vertex normals) or empty
    DevBuf fnrm;      // F*n int32 or empty
    DevBuf rec64;     // F*rec64_stride
    DevBuf build64;   // F*build64_stride (freed after the tree build)
    DevBuf normal32;  // F*n float (geometric normal, flipped per the normal classification)
    DevBuf rec32;     // F*rec32_stride float: normal, d, barycentric planes (simplex-ordered AoS)
    DevBuf ntype;     // F int32: bits 0-1 normal type (0 none, 1 use, 2 reverse), bits 8.. reverse mask
    TreeDev tree;
    int layout = 1;  // FP32 candidate records read by traversal: 0 leaf-ordered SoA, 1 simplex-ordered AoS
    double build_ms = 0;
    double basis_ms = 0;
};

struct Context;

}  // namespace nsp
