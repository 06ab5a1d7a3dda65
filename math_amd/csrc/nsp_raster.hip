// nsp_raster.hip -- rasterised primary visibility for parallel projections
// (gfx950).  See nsp_raster.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <tuple>
#include <vector>

#include "nsp_device.hpp"
#include "nsp_raster.hpp"
#include "nsp_trace.hpp"

namespace nsp {

namespace {

constexpr int RASTER_BLOCK = 256;
constexpr int64_t ITEMS_PER_UNIT = 4096;  // (pixel, sample) pairs one wave takes at a time

struct RasterDims {
    int screen[NSP_MAX_DIM - 1];
    int64_t plane;
    int last;
    int hmax;
    int64_t ppu;
    float dir[NSP_MAX_DIM];  // the projector's direction: every primary ray's dir
};

// Pixel box of each simplex: the projection u_i = (x - cam).a_i / |a_i|^2 +
// size_i / 2 of its vertices (the inverse of ParallelProjector::ray's
// org = cam + sum_i a_i (p_i + jitter_i - size_i / 2)), widened by
// BOX_MARGIN pixels on each side and clipped to the screen.  The margin
// covers the FP32 rounding of the sample origins (about 1e-4 pixel at the
// configurations' scales) with two orders of magnitude to spare; a hit the
// FP32 test accepts lies inside the simplex by more than the re-test band.
constexpr double BOX_MARGIN = 0.01;
template <int N>
__global__ void __launch_bounds__(256) k_raster_box(int64_t F, const double* __restrict__ vtx,
                                                    const int32_t* __restrict__ fac, SceneK<double> kd,
                                                    int32_t* __restrict__ box)
{
    constexpr int D = N - 1;
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= F) return;
    double mn[D], mx[D];
#pragma unroll
    for (int i = 0; i < D; ++i) {
        mn[i] = INFINITY;
        mx[i] = -INFINITY;
    }
    for (int v = 0; v < N; ++v) {
        const double* x = vtx + (int64_t)fac[s * N + v] * N;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            double num = 0, den = 0;
#pragma unroll
            for (int j = 0; j < N; ++j) {
                num += (x[j] - kd.cam_org[j]) * kd.axes[i][j];
                den += kd.axes[i][j] * kd.axes[i][j];
            }
            const double u = num / den + 0.5 * kd.screen[i];
            mn[i] = fmin(mn[i], u);
            mx[i] = fmax(mx[i], u);
        }
    }
    bool empty = false;
    int32_t* b = box + s * 2 * D;
#pragma unroll
    for (int i = 0; i < D; ++i) {
        if (!(mx[i] >= -BOX_MARGIN) || !(mn[i] <= kd.screen[i] + BOX_MARGIN)) empty = true;
        const double lo = fmax(floor(mn[i] - BOX_MARGIN), 0.0);
        const double hi = fmin(floor(mx[i] + BOX_MARGIN), kd.screen[i] - 1.0);
        b[i] = empty ? 1 : (int32_t)lo;
        b[D + i] = empty ? 0 : (int32_t)hi;
        if (!(lo <= hi)) {
            b[i] = 1;
            b[D + i] = 0;
        }
    }
    if (empty) {
        b[0] = 1;
        b[D] = 0;
    }
}

__global__ void k_raster_clear(const BatchDesc* __restrict__ desc, unsigned long long* __restrict__ keys)
{
    const int64_t paths = desc->paths;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < paths; p += (int64_t)gridDim.x * blockDim.x)
        keys[p] = ~0ull;
}

// x / e and x % e for 0 <= x < 2^22 and 1 <= e, with inv = 1 / e in FP32: the
// product x * inv is within 2^-21 x / e < 2 / e of x / e (exact for e = 1), so
// its truncation is within one of the quotient and one correction makes it exact.
__device__ __forceinline__ int small_divmod(int x, int e, float inv, int* rem)
{
    int q = (int)((float)x * inv);
    int r = x - q * e;
    if (r < 0) {
        --q;
        r += e;
    } else if (r >= e) {
        ++q;
        r -= e;
    }
    *rem = r;
    return q;
}

// One wave per work unit (a chunk of one simplex's box): the record is loaded
// once and tested against every sample of the chunk that belongs to the batch.
// SMALL (every box volume and unit item count below 2^22): lane items advance
// the (pixel, sample) pair by 64 without a division, and the box coordinates
// come from float quotients; otherwise 64-bit integer division.
template <int N, bool SMALL>
__global__ void __launch_bounds__(RASTER_BLOCK) k_raster(TreeArgs A, const int32_t* __restrict__ box,
                                                         const int2* __restrict__ unit,
                                                         const int64_t* __restrict__ ustart, RasterDims rd,
                                                         const BatchDesc* __restrict__ desc, int S,
                                                         RayQueue<float> q, int64_t cap,
                                                         unsigned long long* __restrict__ keys)
{
    constexpr int D = N - 1;
    constexpr int RS = rec32_stride(N);
    const int64_t pixel0 = desc->pixel0;
    const int64_t npix = desc->paths / S;
    if (npix <= 0) return;
    const int64_t xlo = pixel0 / rd.plane, xhi = (pixel0 + npix - 1) / rd.plane;
    const int64_t u0 = ustart[std::max<int64_t>(0, xlo - rd.hmax)];
    const int64_t u1 = ustart[std::min<int64_t>(rd.last, xhi + 1)];
    const int lane = (int)(threadIdx.x & 63);
    const int64_t waves = (int64_t)gridDim.x * (RASTER_BLOCK / 64);
    // item it = lane + 64 k is pixel b0 + it / S, sample it % S; 64 = q64 S + r64
    const int q64 = 64 / S, r64 = 64 % S;
    const int lane_q = lane / S, lane_r = lane % S;
    // consecutive units (neighbouring boxes) on one XCD, so their samples share its L2
    for (int64_t u = u0 + (int64_t)xcd_block() * (RASTER_BLOCK / 64) + (threadIdx.x >> 6); u < u1; u += waves) {
        const int2 un = unit[u];
        const int sid = un.x;
        int lo[D], ext[D];
        float inv_ext[D];
        int64_t vol = 1;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            lo[i] = box[(int64_t)sid * 2 * D + i];
            ext[i] = box[(int64_t)sid * 2 * D + D + i] - lo[i] + 1;
            inv_ext[i] = 1.0f / (float)ext[i];
            vol *= ext[i];
        }
        float rec[RS];
        const float4* r4 = reinterpret_cast<const float4*>(A.rec32 + (int64_t)sid * RS);
#pragma unroll
        for (int k = 0; k < RS / 4; ++k) {
            const float4 v = r4[k];
            rec[4 * k] = v.x;
            rec[4 * k + 1] = v.y;
            rec[4 * k + 2] = v.z;
            rec[4 * k + 3] = v.w;
        }
        auto comp = [&](int c) -> float { return rec[c]; };
        auto get_sid = [&]() -> int { return sid; };
        const int64_t b0 = (int64_t)un.y * rd.ppu, b1 = std::min<int64_t>(vol, b0 + rd.ppu);
        const int64_t items = (b1 - b0) * S;
        int next_q = lane_q, next_r = lane_r;
        for (int64_t it = lane; it < items; it += 64) {
            int64_t pixel = 0;
            int s;
            if constexpr (SMALL) {
                int rem = (int)b0 + next_q;
                s = next_r;
                next_r += r64;
                next_q += q64;
                if (next_r >= S) {
                    next_r -= S;
                    ++next_q;
                }
                int64_t stride = 1;
#pragma unroll
                for (int i = 0; i < D - 1; ++i) {
                    int xi;
                    rem = small_divmod(rem, ext[i], inv_ext[i], &xi);
                    pixel += (int64_t)(lo[i] + xi) * stride;
                    stride *= rd.screen[i];
                }
                pixel += (int64_t)(lo[D - 1] + rem) * stride;
            } else {
                const int64_t b = b0 + it / S;
                s = (int)(it % S);
                int64_t rem = b, stride = 1;
#pragma unroll
                for (int i = 0; i < D; ++i) {
                    const int64_t xi = rem % ext[i];
                    rem /= ext[i];
                    pixel += (lo[i] + xi) * stride;
                    stride *= rd.screen[i];
                }
            }
            if (pixel < pixel0 || pixel >= pixel0 + npix) continue;
            const int64_t path = (pixel - pixel0) * S + s;
            float o[N], d[N];
#pragma unroll
            for (int k = 0; k < N; ++k) {
                o[k] = q.org[(int64_t)k * cap + path];
                d[k] = rd.dir[k];
            }
            double th;
            int si;
            bool nr = false;
            unsigned nret = 0;
            if (fp32_test<N, false>(A, comp, get_sid, o, d, nullptr, nullptr, INFINITY, -1, &th, &si, &nr, &nret) ==
                CAND_HIT) {
                const float tf = (float)th;
                const unsigned long long key = ((unsigned long long)__float_as_uint(tf) << 32) | (unsigned)sid;
                atomicMin(&keys[path], key);
            }
        }
    }
}

__global__ void k_raster_resolve(const BatchDesc* __restrict__ desc, RayQueue<float> q,
                                 const unsigned long long* __restrict__ keys)
{
    const int64_t paths = desc->paths;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < paths;
         p += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long key = keys[p];
        const bool hit = key != ~0ull;
        q.hit[p] = hit ? (int32_t)(unsigned)(key & 0xffffffffull) : -1;
        q.hit_t[p] = hit ? __uint_as_float((unsigned)(key >> 32)) : 0.0f;
    }
}

template <int N>
void prepare_t(RasterPrep& rp, const MeshDev& m, const SceneK<double>& kd, int S, hipStream_t s)
{
    constexpr int D = N - 1;
    const int64_t F = m.F;
    rp.box.alloc(sizeof(int32_t) * 2 * D * std::max<int64_t>(F, 1));
    if (F > 0) {
        hipLaunchKernelGGL(k_raster_box<N>, dim3((unsigned)((F + 255) / 256)), dim3(256), 0, s, F,
                           m.vtx64.as<double>(), m.fac.as<int32_t>(), kd, rp.box.as<int32_t>());
        NSP_HIP(hipGetLastError());
    }
    std::vector<int32_t> hb((size_t)2 * D * F);
    NSP_HIP(hipMemcpyAsync(hb.data(), rp.box.p, sizeof(int32_t) * hb.size(), hipMemcpyDeviceToHost, s));
    NSP_HIP(hipStreamSynchronize(s));
    rp.ppu = std::max<int64_t>(1, ITEMS_PER_UNIT / S);
    rp.last = kd.screen[D - 1];
    rp.plane = 1;
    for (int i = 0; i < D; ++i) rp.screen[i] = kd.screen[i];
    for (int i = 0; i < D - 1; ++i) rp.plane *= kd.screen[i];
    // units ordered by the box's low corner, last screen axis major (so that a
    // batch, a range of the last axis, is a contiguous range of units, and
    // consecutive units cover neighbouring pixels)
    std::vector<std::tuple<int64_t, int32_t, int32_t>> us;
    int hmax = 0;
    int64_t max_vol = 0;
    for (int64_t f = 0; f < F; ++f) {
        const int32_t* b = hb.data() + f * 2 * D;
        int64_t vol = 1;
        for (int i = 0; i < D; ++i) vol *= std::max(0, b[D + i] - b[i] + 1);
        if (vol == 0) continue;
        max_vol = std::max(max_vol, vol);
        hmax = std::max(hmax, b[2 * D - 1] - b[D - 1]);
        const int64_t chunks = (vol + rp.ppu - 1) / rp.ppu;
        int64_t corner = 0;
        for (int i = D - 1; i >= 0; --i) corner = corner * kd.screen[i] + b[i];
        for (int64_t c = 0; c < chunks; ++c) us.emplace_back(corner, (int32_t)f, (int32_t)c);
    }
    std::sort(us.begin(), us.end());
    std::vector<int2> hu(us.size());
    std::vector<int64_t> start((size_t)rp.last + 1, (int64_t)us.size());
    for (size_t k = us.size(); k-- > 0;) {
        hu[k] = int2{std::get<1>(us[k]), std::get<2>(us[k])};
        start[std::get<0>(us[k]) / rp.plane] = (int64_t)k;
    }
    for (int v = rp.last - 1; v >= 0; --v) start[v] = std::min(start[v], start[v + 1]);
    rp.hmax = hmax;
    rp.units = (int64_t)hu.size();
    rp.small = max_vol < (int64_t(1) << 22) && rp.ppu * S < (int64_t(1) << 22);
    rp.unit.alloc(sizeof(int2) * std::max<size_t>(hu.size(), 1));
    rp.ustart.alloc(sizeof(int64_t) * start.size());
    if (!hu.empty()) NSP_HIP(hipMemcpyAsync(rp.unit.p, hu.data(), sizeof(int2) * hu.size(), hipMemcpyHostToDevice, s));
    NSP_HIP(hipMemcpyAsync(rp.ustart.p, start.data(), sizeof(int64_t) * start.size(), hipMemcpyHostToDevice, s));
    NSP_HIP(hipStreamSynchronize(s));
    rp.n = N;
    rp.S = S;
    rp.F = F;
}

template <int N>
void launch_t(const RasterPrep& rp, const TreeArgs& A, const BatchDesc* desc, RayQueue<float> q, int64_t cap,
              unsigned long long* keys, const float* dir, hipStream_t s)
{
    const unsigned gclear = 4096;
    hipLaunchKernelGGL(k_raster_clear, dim3(gclear), dim3(256), 0, s, desc, keys);
    NSP_HIP(hipGetLastError());
    RasterDims rd{};
    for (int i = 0; i < N - 1; ++i) rd.screen[i] = rp.screen[i];
    rd.plane = rp.plane;
    rd.last = rp.last;
    rd.hmax = rp.hmax;
    rd.ppu = rp.ppu;
    for (int k = 0; k < N; ++k) rd.dir[k] = dir[k];
    const unsigned blocks = 2048;  // 8192 waves; each wave loops over the batch's units
    if (rp.small)
        hipLaunchKernelGGL((k_raster<N, true>), dim3(blocks), dim3(RASTER_BLOCK), 0, s, A, rp.box.as<int32_t>(),
                           rp.unit.as<int2>(), rp.ustart.as<int64_t>(), rd, desc, rp.S, q, cap, keys);
    else
        hipLaunchKernelGGL((k_raster<N, false>), dim3(blocks), dim3(RASTER_BLOCK), 0, s, A, rp.box.as<int32_t>(),
                           rp.unit.as<int2>(), rp.ustart.as<int64_t>(), rd, desc, rp.S, q, cap, keys);
    NSP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_raster_resolve, dim3(gclear), dim3(256), 0, s, desc, q, keys);
    NSP_HIP(hipGetLastError());
}

}  // namespace

void raster_prepare(RasterPrep& rp, const MeshDev& m, const SceneK<double>& kd, int S, hipStream_t s)
{
    if (rp.n == m.n && rp.S == S && rp.F == m.F && rp.unit.p) return;
    switch (m.n) {
    case 3: prepare_t<3>(rp, m, kd, S, s); break;
    case 4: prepare_t<4>(rp, m, kd, S, s); break;
    case 5: prepare_t<5>(rp, m, kd, S, s); break;
    case 6: prepare_t<6>(rp, m, kd, S, s); break;
    default: raise(NSP_E_INVALID, "dimension not supported");
    }
}

void raster_launch(const RasterPrep& rp, const MeshDev& m, const TreeArgs& A, const BatchDesc* desc,
                   RayQueue<float> q, int64_t cap, unsigned long long* keys, const float* dir, hipStream_t s)
{
    switch (m.n) {
    case 3: launch_t<3>(rp, A, desc, q, cap, keys, dir, s); break;
    case 4: launch_t<4>(rp, A, desc, q, cap, keys, dir, s); break;
    case 5: launch_t<5>(rp, A, desc, q, cap, keys, dir, s); break;
    case 6: launch_t<6>(rp, A, desc, q, cap, keys, dir, s); break;
    default: raise(NSP_E_INVALID, "dimension not supported");
    }
}

}  // namespace nsp
