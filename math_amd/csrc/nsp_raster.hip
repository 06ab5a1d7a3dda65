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
#ifndef NSP_ITEMS_PER_UNIT
#define NSP_ITEMS_PER_UNIT 8192
#endif
// (pixel, sample) pairs one wave takes at a time (C1: 8192 items 3.48 ms per
// pass, 4096 3.53, 2048 3.75)
constexpr int64_t ITEMS_PER_UNIT = NSP_ITEMS_PER_UNIT;

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
// Besides the box, the planes of the projected simplex (a (N-1)-simplex of
// the screen): per vertex k, the unit normal (in pixels) of the projected face
// opposite k, pointing towards k, and its offset, so that n.u + c is the
// distance of screen point u from that face's plane, positive inside.  A pixel
// [x, x + 1)^(N-1) whose largest distance over the pixel is below -BOX_MARGIN
// for some face holds no sample inside the simplex dilated by BOX_MARGIN, the
// same margin the box keeps.  When the projection is degenerate (a face or the
// simplex of near-zero extent on the screen), the offsets are +inf and only the
// box limits the samples.
template <int N>
__global__ void __launch_bounds__(256) k_raster_box(int64_t F, const double* __restrict__ vtx,
                                                    const int32_t* __restrict__ fac, SceneK<double> kd,
                                                    int32_t* __restrict__ box, float* __restrict__ planes)
{
    constexpr int D = N - 1;
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= F) return;
    double P[N][D];
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
            P[v][i] = u;
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
    // face planes; scale: the box diagonal, for the degeneracy thresholds
    double scale = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) scale += (mx[i] - mn[i]) * (mx[i] - mn[i]);
    scale = sqrt(scale);
    double nrm[N][D], off[N];
    bool degenerate = !(scale > 0);
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const int j0 = k == 0 ? 1 : 0;
        double m[(D > 1 ? D - 1 : 1) * D];
        int r = 0;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            if (j == k || j == j0) continue;
#pragma unroll
            for (int i = 0; i < D; ++i) m[r * D + i] = P[j][i] - P[j0][i];
            ++r;
        }
        double nk[D];
        ortho_nn<D, double>(m, nk);
        double len = 0, h = 0;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            len += nk[i] * nk[i];
            h += nk[i] * (P[k][i] - P[j0][i]);
        }
        len = sqrt(len);
        h = h / len;
        if (!(len > 1e-12 * pow(scale, (double)(D - 1))) || !(fabs(h) > 1e-9 * scale)) degenerate = true;
        const double sg = h < 0 ? -1.0 / len : 1.0 / len;
        double c = 0;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            nrm[k][i] = nk[i] * sg;
            c -= nrm[k][i] * P[j0][i];
        }
        off[k] = c;
    }
    float* pl = planes + s * N * N;
#pragma unroll
    for (int k = 0; k < N; ++k) {
#pragma unroll
        for (int i = 0; i < D; ++i) pl[k * N + i] = degenerate ? 0.0f : (float)nrm[k][i];
        pl[k * N + D] = degenerate ? INFINITY : (float)off[k];
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

// One wave per work unit (a chunk of one simplex's box): the record and face
// planes are loaded once.  The wave takes the chunk's pixels 64 at a time: each
// lane decodes one pixel and keeps it if it belongs to the batch and is not
// culled by a face plane; the kept pixels are packed to the low lanes
// (ds_permute), and then every (kept pixel, sample) pair gets the FP32 test.
// SMALL (every box volume below 2^22): the box coordinates come from float
// quotients; otherwise 64-bit integer division.
template <int N, bool SMALL>
__global__ void __launch_bounds__(RASTER_BLOCK) k_raster(TreeArgs A, const int32_t* __restrict__ box,
                                                         const float* __restrict__ planes,
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
    const float inv_s = 1.0f / (float)S;
    // (pair index) / S by a float quotient while pair indices stay below 2^22
    const bool small_s = (int64_t)S * 128 < (int64_t(1) << 22);
    const float margin = (float)BOX_MARGIN;
    // consecutive units (neighbouring boxes) on one XCD, so their samples share its L2
    // the unit is the wave's: its simplex's box, planes and record are read with
    // scalar loads, and the next unit's index is loaded during this one
    int64_t u = u0 + (int64_t)xcd_block() * (RASTER_BLOCK / 64) + (threadIdx.x >> 6);
    int2 un_next = u < u1 ? unit[u] : int2{0, 0};
    for (; u < u1; u += waves) {
        const int sid = __builtin_amdgcn_readfirstlane(un_next.x);
        const int chunk = __builtin_amdgcn_readfirstlane(un_next.y);
        if (u + waves < u1) un_next = unit[u + waves];
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
        float pl[N * N];
#pragma unroll
        for (int k = 0; k < N * N; ++k) pl[k] = planes[(int64_t)sid * N * N + k];
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
        const int64_t b0 = (int64_t)chunk * rd.ppu, b1 = std::min<int64_t>(vol, b0 + rd.ppu);
        for (int64_t g = b0; g < b1; g += 64) {
            // this lane's pixel of the group
            const int64_t pb = g + lane;
            bool keep = pb < b1;
            int rel = 0;
            if (keep) {
                int x[D];
                if constexpr (SMALL) {
                    int rem = (int)pb;
#pragma unroll
                    for (int i = 0; i < D - 1; ++i) rem = small_divmod(rem, ext[i], inv_ext[i], &x[i]);
                    x[D - 1] = rem;
                } else {
                    int64_t rem = pb;
#pragma unroll
                    for (int i = 0; i < D; ++i) {
                        x[i] = (int)(rem % ext[i]);
                        rem /= ext[i];
                    }
                }
                int64_t pixel = 0, stride = 1;
#pragma unroll
                for (int i = 0; i < D; ++i) {
                    x[i] += lo[i];
                    pixel += (int64_t)x[i] * stride;
                    stride *= rd.screen[i];
                }
                keep = pixel >= pixel0 && pixel < pixel0 + npix;
                rel = (int)(pixel - pixel0);
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    // the largest distance from face k's plane over the pixel
                    float dmax = pl[k * N + D];
#pragma unroll
                    for (int i = 0; i < D; ++i) {
                        const float w = pl[k * N + i];
                        dmax += w * (float)(x[i] + (w > 0.0f ? 1 : 0));
                    }
                    keep = keep && !(dmax < -margin);
                }
            }
            const uint64_t mask = __ballot(keep);
            if (mask == 0) continue;
            const int cnt = __popcll(mask);
            const int below = __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0));
            // kept pixels to lanes 0..cnt-1 in lane order, the others after them
            const int dst = keep ? below : cnt + (lane - below);
            const int packed = __builtin_amdgcn_ds_permute(dst << 2, rel);
            const int items = cnt * S;
            for (int base = 0; base < items; base += 64) {
                const int it = base + lane;
                int s, r;
                if (small_s) {
                    r = small_divmod(it, S, inv_s, &s);
                } else {
                    r = it / S;
                    s = it - r * S;
                }
                const int prel = __builtin_amdgcn_ds_bpermute((r < 64 ? r : 63) << 2, packed);
                if (it >= items) continue;
                const int64_t path = (int64_t)prel * S + s;
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
                if (fp32_test<N, false>(A, comp, get_sid, o, d, nullptr, nullptr, INFINITY, -1, &th, &si, &nr,
                                        &nret) == CAND_HIT) {
                    const float tf = (float)th;
                    const unsigned long long key = ((unsigned long long)__float_as_uint(tf) << 32) | (unsigned)sid;
                    atomicMin(&keys[path], key);
                }
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
    rp.planes.alloc(sizeof(float) * N * N * std::max<int64_t>(F, 1));
    if (F > 0) {
        hipLaunchKernelGGL(k_raster_box<N>, dim3((unsigned)((F + 255) / 256)), dim3(256), 0, s, F,
                           m.vtx64.as<double>(), m.fac.as<int32_t>(), kd, rp.box.as<int32_t>(),
                           rp.planes.as<float>());
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
    rp.small = max_vol < (int64_t(1) << 22);
    if ((int64_t)S * 128 > INT32_MAX) raise(NSP_E_UNSUPPORTED, "too many samples per pixel for rasterising");
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
#ifndef NSP_RASTER_BLOCKS
#define NSP_RASTER_BLOCKS 4096
#endif
    // 16384 waves, twice what fits at once: each wave loops over the batch's units,
    // and the second round of blocks evens out the waves' shares (C1: 4096 blocks
    // 3.08 ms, 2048 3.47)
    const unsigned blocks = NSP_RASTER_BLOCKS;
    if (rp.small)
        hipLaunchKernelGGL((k_raster<N, true>), dim3(blocks), dim3(RASTER_BLOCK), 0, s, A, rp.box.as<int32_t>(),
                           rp.planes.as<float>(), rp.unit.as<int2>(), rp.ustart.as<int64_t>(), rd, desc, rp.S, q,
                           cap, keys);
    else
        hipLaunchKernelGGL((k_raster<N, false>), dim3(blocks), dim3(RASTER_BLOCK), 0, s, A, rp.box.as<int32_t>(),
                           rp.planes.as<float>(), rp.unit.as<int2>(), rp.ustart.as<int64_t>(), rd, desc, rp.S, q,
                           cap, keys);
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
