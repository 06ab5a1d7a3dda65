// nsp_shadow_raster.hip -- shadow rays towards one point light by rasterising
// the simplices as the light sees them (gfx950).  See nsp_raster.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "nsp_device.hpp"
#include "nsp_raster.hpp"
#include "nsp_stages.hpp"
#include "nsp_trace.hpp"

namespace nsp {

namespace {

constexpr int SR_BLOCK = 256;
// Margin, in cells, by which footprints are widened.  A shadow ray o + t d runs
// from its origin (its hit point moved by the ray offset towards the light) to
// the light; its FP32 direction passes the light within about 1e-7 of its
// length, so where it meets a simplex its projection lies within about 1e-7
// (distance / depth) of its hit point's, far below the margin at the cell sizes
// used (at least 1e-4 of the projected extent).
constexpr double SR_MARGIN = 0.01;
#ifndef NSP_SR_RASTER_BLOCKS
#define NSP_SR_RASTER_BLOCKS 4096
#endif
constexpr int SR_CELLS_PER_UNIT = 64;
constexpr int SR_SCAN_TILE = 4096;  // cells per block of the count scan (16 per thread)
// Cell edge: the mean footprint extent per axis over SR_CELL_DIV.  Every ray is
// tested against each simplex whose widened footprint meets its cell, so
// smaller cells mean fewer tests per ray and more cells per footprint.
#ifndef NSP_SR_CELL_DIV
#define NSP_SR_CELL_DIV 4
#endif
constexpr double SR_CELL_DIV = NSP_SR_CELL_DIV;
constexpr int SR_MAX_TILES = 4096;  // at most 2^24 cells

// Floats per binned shadow ray: its hit point (the ray is rebuilt from it with
// shadow_ray_from, as shading built it).
__host__ __device__ constexpr int sr_record(int n) { return n; }

// The light as shading saw it: FP32 position and the ray offset.
struct SRLight {
    float pos[NSP_MAX_DIM];
    float offset;
};

// Screen coordinates of a point x as the light sees it: s_i = e_i.(x - light) /
// a.(x - light), a the axis from the light to the mesh centre, e_i an
// orthonormal basis of its complement (a central projection onto the plane at
// distance 1); cell coordinates u_i = (s_i - smin_i) * inv_cell_i.
template <int N>
__device__ __forceinline__ bool light_coords(const SRGrid& g, const double* x, double* u)
{
    double r[N];
#pragma unroll
    for (int k = 0; k < N; ++k) r[k] = x[k] - g.light[k];
    double z = 0;
#pragma unroll
    for (int k = 0; k < N; ++k) z += g.axis[k] * r[k];
#pragma unroll
    for (int i = 0; i < N - 1; ++i) {
        double s = 0;
#pragma unroll
        for (int k = 0; k < N; ++k) s += g.basis[i][k] * r[k];
        u[i] = (s / z - g.smin[i]) * g.inv_cell[i];
    }
    return z > 0;
}

// Per vertex: its depth along the axis and cell coordinates.
template <int N>
__global__ void __launch_bounds__(256) k_sr_vertices(int64_t V, const double* __restrict__ vtx, SRGrid g,
                                                     double* __restrict__ out)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    double r[N];
#pragma unroll
    for (int k = 0; k < N; ++k) r[k] = vtx[v * N + k] - g.light[k];
    double z = 0;
#pragma unroll
    for (int k = 0; k < N; ++k) z += g.axis[k] * r[k];
    out[v * N] = z;
#pragma unroll
    for (int i = 0; i < N - 1; ++i) {
        double s = 0;
#pragma unroll
        for (int k = 0; k < N; ++k) s += g.basis[i][k] * r[k];
        out[v * N + 1 + i] = (s / z - g.smin[i]) * g.inv_cell[i];
    }
}

// Face planes of each simplex's footprint in cell coordinates (as the primary
// rasteriser's, k_raster_box): per vertex k the unit normal of the projected
// face opposite k, towards k, and its offset; +inf offsets when the footprint
// is degenerate.  The footprint of a simplex all of whose vertices lie in
// front of the light is the simplex of its projected vertices.
template <int N>
__global__ void __launch_bounds__(256) k_sr_planes(int64_t F, const int32_t* __restrict__ fac,
                                                   const double* __restrict__ vu, float* __restrict__ planes)
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
        const int64_t x = fac[s * N + v];
#pragma unroll
        for (int i = 0; i < D; ++i) {
            P[v][i] = vu[x * N + 1 + i];
            mn[i] = fmin(mn[i], P[v][i]);
            mx[i] = fmax(mx[i], P[v][i]);
        }
    }
    double scale = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) scale += (mx[i] - mn[i]) * (mx[i] - mn[i]);
    scale = sqrt(scale);
    bool degenerate = !(scale > 0);
    double nrm[N][D], off[N];
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

// The cell of each shadow ray of the depth (its hit point's), counted per cell.
// Block b takes queue segment b % QSEG.
template <int N>
__global__ void __launch_bounds__(SR_BLOCK) k_sr_count(ShadowQueue<float> sq, int64_t scap,
                                                       const unsigned long long* __restrict__ dc, SRGrid g,
                                                       int32_t* __restrict__ cell_of, uint32_t* __restrict__ rank,
                                                       uint32_t* __restrict__ count)
{
    const int seg = (int)(blockIdx.x % QSEG);
    const int64_t segcap = (int64_t)dc[CNT_SSEGCAP];
    const int64_t n = (int64_t)dc[cnt_shadowseg(seg)];
    const int64_t stride = (int64_t)(gridDim.x / QSEG) * blockDim.x;
    for (int64_t j = (int64_t)(blockIdx.x / QSEG) * blockDim.x + threadIdx.x; j < n; j += stride) {
        const int64_t i = seg * segcap + j;
        double o[N], u[N - 1];
#pragma unroll
        for (int k = 0; k < N; ++k) o[k] = (double)sq.org[(int64_t)k * scap + i];
        light_coords<N>(g, o, u);
        int64_t c = 0;
#pragma unroll
        for (int d = N - 2; d >= 0; --d) {
            const double x = floor(u[d]);
            const int gi = !(x >= 0) ? 0 : x >= g.grid[d] ? g.grid[d] - 1 : (int)x;  // NaN to cell 0
            c = c * g.grid[d] + gi;
        }
        cell_of[i] = (int32_t)c;
        rank[i] = atomicAdd(&count[c], 1u);
    }
}

// Exclusive scan of the per-cell counts in three launches: tile sums, the
// scan of the tile sums (one block), then each tile's prefixes, written to
// start while the counts are cleared for the next depth.
__global__ void __launch_bounds__(256) k_sr_tile_sums(const uint32_t* __restrict__ count, int64_t cells,
                                                      uint32_t* __restrict__ tiles)
{
    __shared__ uint32_t part[256];
    const int64_t t0 = (int64_t)blockIdx.x * SR_SCAN_TILE;
    uint32_t s = 0;
    for (int k = threadIdx.x; k < SR_SCAN_TILE; k += 256) {
        const int64_t c = t0 + k;
        if (c < cells) s += count[c];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) tiles[blockIdx.x] = part[0];
}

__global__ void __launch_bounds__(1024) k_sr_tile_scan(uint32_t* __restrict__ tiles, int ntiles,
                                                       uint32_t* __restrict__ total)
{
    __shared__ uint32_t buf[SR_MAX_TILES];
    for (int k = threadIdx.x; k < ntiles; k += 1024) buf[k] = tiles[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int k = 0; k < ntiles; ++k) {
            const uint32_t v = buf[k];
            buf[k] = run;
            run += v;
        }
        *total = run;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < ntiles; k += 1024) tiles[k] = buf[k];
}

__global__ void __launch_bounds__(256) k_sr_tile_down(uint32_t* __restrict__ count, int64_t cells,
                                                      const uint32_t* __restrict__ tiles,
                                                      const uint32_t* __restrict__ total, uint32_t* __restrict__ start)
{
    // each thread takes 16 consecutive cells; a block scan of the thread sums
    constexpr int PER = SR_SCAN_TILE / 256;
    __shared__ uint32_t part[256];
    const int64_t c0 = (int64_t)blockIdx.x * SR_SCAN_TILE + (int64_t)threadIdx.x * PER;
    uint32_t v[PER];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        v[k] = c0 + k < cells ? count[c0 + k] : 0u;
        s += v[k];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const uint32_t y = (int)threadIdx.x >= o ? part[threadIdx.x - o] : 0u;
        __syncthreads();
        part[threadIdx.x] += y;
        __syncthreads();
    }
    uint32_t run = tiles[blockIdx.x] + part[threadIdx.x] - s;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        if (c0 + k < cells) {
            start[c0 + k] = run;
            count[c0 + k] = 0u;
        }
        run += v[k];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) start[cells] = *total;
}

// Each shadow ray's hit point to its place in cell order.
template <int N>
__global__ void __launch_bounds__(SR_BLOCK) k_sr_scatter(ShadowQueue<float> sq, int64_t scap,
                                                         const unsigned long long* __restrict__ dc,
                                                         const int32_t* __restrict__ cell_of,
                                                         const uint32_t* __restrict__ rank,
                                                         const uint32_t* __restrict__ start, float* __restrict__ rays,
                                                         int32_t* __restrict__ ray_of, uint8_t* __restrict__ occl_q)
{
    const int seg = (int)(blockIdx.x % QSEG);
    const int64_t segcap = (int64_t)dc[CNT_SSEGCAP];
    const int64_t n = (int64_t)dc[cnt_shadowseg(seg)];
    const int64_t stride = (int64_t)(gridDim.x / QSEG) * blockDim.x;
    for (int64_t j = (int64_t)(blockIdx.x / QSEG) * blockDim.x + threadIdx.x; j < n; j += stride) {
        const int64_t i = seg * segcap + j;
        const int64_t pos = (int64_t)start[cell_of[i]] + rank[i];
        constexpr int R = sr_record(N);
        float r[R];
#pragma unroll
        for (int k = 0; k < N; ++k) r[k] = sq.org[(int64_t)k * scap + i];
        float* dst = rays + pos * R;
#pragma unroll
        for (int k = 0; k < R; ++k) dst[k] = r[k];
        ray_of[pos] = (int32_t)i;
        occl_q[i] = 0;
    }
}

// One wave per work unit (up to 64 cells of one simplex's footprint box): the
// record and face planes are loaded once, the unit's cells culled by the
// planes, and every shadow ray of the kept cells gets the any-hit test the
// shadow walk applies to a leaf candidate (FP32 with the FP64 re-test inline).
// The rays of the kept cells are dealt 64 at a time: a wave scan of the cells'
// ray counts, and a lane finds its cell by a binary search over the scan.  A
// hit writes a NaN into the ray's record, so the record load that brings a ray
// to a later simplex also tells it to skip the test.
template <int N>
__global__ void __launch_bounds__(SR_BLOCK) k_sr_raster(TreeArgs A, const int2* __restrict__ unit, int64_t units,
                                                        const int32_t* __restrict__ box,
                                                        const float* __restrict__ planes, SRGrid g,
                                                        const uint32_t* __restrict__ start,
                                                        float* __restrict__ rays, int64_t scap,
                                                        const int32_t* __restrict__ ray_of,
                                                        uint8_t* __restrict__ occl_q, SRLight lt)
{
    constexpr int D = N - 1;
    constexpr int RS = rec32_stride(N);
    const int lane = (int)(threadIdx.x & 63);
    const int64_t waves = (int64_t)gridDim.x * (SR_BLOCK / 64);
    const float margin = (float)SR_MARGIN;
    // the unit is the wave's: its simplex's box, planes and record are read with
    // scalar loads, and the next unit's index is loaded during this one
    int64_t u = (int64_t)xcd_block() * (SR_BLOCK / 64) + (threadIdx.x >> 6);
    int2 un_next = u < units ? unit[u] : int2{0, 0};
    for (; u < units; u += waves) {
        const int sid = __builtin_amdgcn_readfirstlane(un_next.x);
        const int chunk = __builtin_amdgcn_readfirstlane(un_next.y);
        if (u + waves < units) un_next = unit[u + waves];
        int lo[D], ext[D];
        int vol = 1;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            lo[i] = box[(int64_t)sid * 2 * D + i];
            ext[i] = box[(int64_t)sid * 2 * D + D + i] - lo[i] + 1;
            vol *= ext[i];
        }
        float pl[N * N];
#pragma unroll
        for (int k = 0; k < N * N; ++k) pl[k] = planes[(int64_t)sid * N * N + k];
        float rec[RS];
        load_rec32<N>(A, sid, rec);
        const int pb = chunk * SR_CELLS_PER_UNIT + lane;
        bool keep = pb < vol;
        uint32_t first = 0, cnt = 0;
        if (keep) {
            int x[D];
            int rem = pb;
#pragma unroll
            for (int i = 0; i < D - 1; ++i) {
                const int q = rem / ext[i];
                x[i] = rem - q * ext[i] + lo[i];
                rem = q;
            }
            x[D - 1] = rem + lo[D - 1];
#pragma unroll
            for (int k = 0; k < N; ++k) {
                float dmax = pl[k * N + D];
#pragma unroll
                for (int i = 0; i < D; ++i) {
                    const float w = pl[k * N + i];
                    dmax += w * (float)(x[i] + (w > 0.0f ? 1 : 0));
                }
                keep = keep && !(dmax < -margin);
            }
            if (keep) {
                int64_t c = 0;
#pragma unroll
                for (int i = D - 1; i >= 0; --i) c = c * g.grid[i] + x[i];
                first = start[c];
                cnt = start[c + 1] - first;
            }
        }
        if (__ballot(cnt > 0) == 0) continue;
        uint32_t incl = cnt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(incl, o);
            if (lane >= o) incl += y;
        }
        const uint32_t total = __shfl(incl, 63);
        // a ray's place in cell order: its slot j plus its cell's first ray minus the rays before its cell
        const int64_t delta = (int64_t)first - (int64_t)(incl - cnt);
        auto comp = [&](int cc) -> float { return rec[cc]; };
        auto get_sid = [&]() -> int { return sid; };
        for (uint32_t base = 0; base < total; base += 64) {
            const uint32_t j = base + (uint32_t)lane;
            // owner: the number of lanes whose inclusive count is at most j
            int owner = 0;
#pragma unroll
            for (int step = 32; step > 0; step >>= 1) {
                const uint32_t v = __shfl(incl, owner + step - 1);
                if (v <= j) owner += step;
            }
            const int ow = owner < 64 ? owner : 63;
            const int64_t pos = (int64_t)j + __shfl(delta, ow);
            if (j >= total) continue;
            constexpr int R = sr_record(N);
            float r[R];
            const float* src = rays + pos * R;
#pragma unroll
            for (int k = 0; k < R; ++k) r[k] = src[k];
            // an occluded ray's record holds a NaN (a hit point is finite)
            if (!(r[0] == r[0])) continue;
            float vl[N], o[N], d[N], tmax;
            shadow_ray_from<N, float>(lt.pos, r, lt.offset, vl, d, &tmax, o);
            double th;
            int si;
            bool nr = false;
            unsigned nret = 0;
            if (fp32_test<N, false>(A, comp, get_sid, o, d, nullptr, nullptr, tmax, -1, &th, &si, &nr, &nret) ==
                CAND_HIT) {
                rays[pos * R] = __int_as_float(0x7fc00000);
                occl_q[ray_of[pos]] = 1;
            }
        }
    }
}

// The contributions of the rays that reached the light, in queue order (the
// queue's paths, contributions and the paths' radiance are read in order).
// Block b takes queue segment b % QSEG.
template <int N>
__global__ void __launch_bounds__(SR_BLOCK) k_sr_finish(ShadowQueue<float> sq, int64_t scap,
                                                        const unsigned long long* __restrict__ dc, int64_t cells,
                                                        const uint32_t* __restrict__ start,
                                                        const uint8_t* __restrict__ occl_q, float* __restrict__ Ls,
                                                        int64_t cap, float* __restrict__ direct,
                                                        unsigned long long* __restrict__ rays)
{
    if (rays && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(rays, (unsigned long long)start[cells]);
    const int seg = (int)(blockIdx.x % QSEG);
    const int64_t segcap = (int64_t)dc[CNT_SSEGCAP];
    const int64_t n = (int64_t)dc[cnt_shadowseg(seg)];
    const int64_t stride = (int64_t)(gridDim.x / QSEG) * blockDim.x;
    for (int64_t j = (int64_t)(blockIdx.x / QSEG) * blockDim.x + threadIdx.x; j < n; j += stride) {
        const int64_t i = seg * segcap + j;
        if (!occl_q[i]) shadow_unoccluded<N, float>(sq, scap, i, Ls, cap, direct);
    }
}

template <int N>
void prepare_t(ShadowRasterPrep& sp, const MeshDev& m, const double* light, hipStream_t s)
{
    constexpr int D = N - 1;
    SRGrid& g = sp.grid;
    g = SRGrid{};
    sp.ok = false;
    double axis[N], len = 0;
    for (int k = 0; k < N; ++k) {
        g.light[k] = light[k];
        axis[k] = 0.5 * (m.mn[k] + m.mx[k]) - light[k];
        len += axis[k] * axis[k];
    }
    len = std::sqrt(len);
    if (!(len > 0)) return;
    for (int k = 0; k < N; ++k) g.axis[k] = axis[k] / len;
    // orthonormal complement of the axis: Gram-Schmidt over the coordinate axes
    int made = 0;
    for (int e = 0; e < N && made < D; ++e) {
        double v[N];
        for (int k = 0; k < N; ++k) v[k] = (k == e) ? 1.0 : 0.0;
        for (int pass = 0; pass < 2; ++pass) {
            double a = 0;
            for (int k = 0; k < N; ++k) a += v[k] * g.axis[k];
            for (int k = 0; k < N; ++k) v[k] -= a * g.axis[k];
            for (int q = 0; q < made; ++q) {
                double b = 0;
                for (int k = 0; k < N; ++k) b += v[k] * g.basis[q][k];
                for (int k = 0; k < N; ++k) v[k] -= b * g.basis[q][k];
            }
        }
        double l = 0;
        for (int k = 0; k < N; ++k) l += v[k] * v[k];
        l = std::sqrt(l);
        if (l < 0.3) continue;
        for (int k = 0; k < N; ++k) g.basis[made][k] = v[k] / l;
        ++made;
    }
    if (made != D) return;
    // vertex depths and projections (smin = 0, inv_cell = 1: screen coordinates)
    for (int i = 0; i < D; ++i) {
        g.smin[i] = 0;
        g.inv_cell[i] = 1;
    }
    const int64_t V = m.V, F = m.F;
    if (V <= 0 || F <= 0) return;
    DevBuf vu;
    vu.alloc(sizeof(double) * N * V);
    hipLaunchKernelGGL(k_sr_vertices<N>, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, s, V,
                       m.vtx64.as<double>(), g, vu.as<double>());
    NSP_HIP(hipGetLastError());
    std::vector<double> hv((size_t)N * V);
    std::vector<int32_t> hf((size_t)N * F);
    NSP_HIP(hipMemcpyAsync(hv.data(), vu.p, sizeof(double) * hv.size(), hipMemcpyDeviceToHost, s));
    NSP_HIP(hipMemcpyAsync(hf.data(), m.fac.p, sizeof(int32_t) * hf.size(), hipMemcpyDeviceToHost, s));
    NSP_HIP(hipStreamSynchronize(s));
    // every vertex well in front of the light (so that the footprints are the
    // projected simplices and every shadow ray's hit point projects)
    double dist = 0;
    for (int k = 0; k < N; ++k) dist = std::max(dist, std::max(std::fabs(m.mn[k] - light[k]), std::fabs(m.mx[k] - light[k])));
    double zmin = INFINITY;
    double smin[D], smax[D], ext_sum[D];
    for (int i = 0; i < D; ++i) {
        smin[i] = INFINITY;
        smax[i] = -INFINITY;
        ext_sum[i] = 0;
    }
    std::vector<char> used((size_t)V, 0);
    for (int64_t f = 0; f < F; ++f)
        for (int v = 0; v < N; ++v) used[hf[(size_t)f * N + v]] = 1;
    for (int64_t v = 0; v < V; ++v) {
        if (!used[v]) continue;
        zmin = std::min(zmin, hv[(size_t)v * N]);
        for (int i = 0; i < D; ++i) {
            const double x = hv[(size_t)v * N + 1 + i];
            smin[i] = std::min(smin[i], x);
            smax[i] = std::max(smax[i], x);
        }
    }
    if (!(zmin > 1e-2 * dist) || !std::isfinite(zmin)) return;
    // cell size: half the mean footprint extent per axis, at least 1/4096 of
    // the projected extent, at most 2^24 cells
    for (int64_t f = 0; f < F; ++f)
        for (int i = 0; i < D; ++i) {
            double a = INFINITY, b = -INFINITY;
            for (int v = 0; v < N; ++v) {
                const double x = hv[(size_t)hf[(size_t)f * N + v] * N + 1 + i];
                a = std::min(a, x);
                b = std::max(b, x);
            }
            ext_sum[i] += b - a;
        }
    // the largest drift of a shadow ray's projection from its origin's (FP32
    // rounding of its direction and of the light position it aims at, about
    // 1e-7 of the coordinates, over the nearest depth), kept below a tenth of
    // the margin
    double amax = 0;
    for (int k = 0; k < N; ++k)
        amax = std::max(amax, std::max(std::fabs(light[k]), std::max(std::fabs(m.mn[k]), std::fabs(m.mx[k]))));
    const double drift = 1e-6 * amax / zmin;
    double cell[D];
    double cells_total = 1;
    for (int i = 0; i < D; ++i) {
        const double span = smax[i] - smin[i];
        if (!(span > 0)) return;
        cell[i] = std::max(std::max(ext_sum[i] / (double)F / SR_CELL_DIV, span / 4096.0), drift / SR_MARGIN);
        cells_total *= std::ceil(span / cell[i]) + 1;
    }
    const double limit = (double)SR_SCAN_TILE * SR_MAX_TILES;
    if (cells_total > limit) {
        const double grow = std::pow(cells_total / limit, 1.0 / D) * 1.01;
        for (int i = 0; i < D; ++i) cell[i] *= grow;
    }
    int64_t cells = 1;
    for (int i = 0; i < D; ++i) {
        g.smin[i] = smin[i];
        g.inv_cell[i] = 1.0 / cell[i];
        g.grid[i] = (int)std::ceil((smax[i] - smin[i]) / cell[i]) + 1;
        cells *= g.grid[i];
    }
    if (cells > (int64_t)limit) return;
    sp.cells = cells;
    // cell coordinates of the vertices, footprint boxes and work units (in
    // simplex order; neighbouring simplices cover neighbouring cells)
    hipLaunchKernelGGL(k_sr_vertices<N>, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, s, V,
                       m.vtx64.as<double>(), g, vu.as<double>());
    NSP_HIP(hipGetLastError());
    NSP_HIP(hipMemcpyAsync(hv.data(), vu.p, sizeof(double) * hv.size(), hipMemcpyDeviceToHost, s));
    NSP_HIP(hipStreamSynchronize(s));
    std::vector<int32_t> hb((size_t)2 * D * F);
    std::vector<int2> hu;
    hu.reserve((size_t)F * 2);
    for (int64_t f = 0; f < F; ++f) {
        int64_t vol = 1;
        for (int i = 0; i < D; ++i) {
            double a = INFINITY, b = -INFINITY;
            for (int v = 0; v < N; ++v) {
                const double x = hv[(size_t)hf[(size_t)f * N + v] * N + 1 + i];
                a = std::min(a, x);
                b = std::max(b, x);
            }
            const int lo = (int)std::max(0.0, std::floor(a - SR_MARGIN));
            const int hi = (int)std::min((double)g.grid[i] - 1, std::floor(b + SR_MARGIN));
            hb[(size_t)f * 2 * D + i] = lo;
            hb[(size_t)f * 2 * D + D + i] = hi;
            vol *= std::max(0, hi - lo + 1);
        }
        if (vol > INT32_MAX / 2) return;
        const int64_t chunks = (vol + SR_CELLS_PER_UNIT - 1) / SR_CELLS_PER_UNIT;
        for (int64_t c = 0; c < chunks; ++c) hu.push_back(int2{(int)f, (int)c});
    }
    // units in the Morton order of their boxes' low corners, so that the
    // simplices that cover a compact region of cells run together on one XCD
    // and the region's rays stay in that XCD's L2 between them
    {
        std::vector<std::pair<uint64_t, int64_t>> key(hu.size());
        for (size_t k = 0; k < hu.size(); ++k) {
            const int32_t* bx = hb.data() + (size_t)hu[k].x * 2 * D;
            uint64_t code = 0;
            for (int bit = 11; bit >= 0; --bit)
                for (int i = D - 1; i >= 0; --i) code = (code << 1) | (uint64_t)((bx[i] >> bit) & 1);
            key[k] = {code, (int64_t)k};
        }
        std::sort(key.begin(), key.end());
        std::vector<int2> sorted(hu.size());
        for (size_t k = 0; k < hu.size(); ++k) sorted[k] = hu[(size_t)key[k].second];
        hu.swap(sorted);
    }
    sp.units = (int64_t)hu.size();
    sp.box.alloc(sizeof(int32_t) * hb.size());
    sp.unit.alloc(sizeof(int2) * std::max<size_t>(hu.size(), 1));
    sp.planes.alloc(sizeof(float) * N * N * F);
    NSP_HIP(hipMemcpyAsync(sp.box.p, hb.data(), sizeof(int32_t) * hb.size(), hipMemcpyHostToDevice, s));
    if (!hu.empty())
        NSP_HIP(hipMemcpyAsync(sp.unit.p, hu.data(), sizeof(int2) * hu.size(), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_sr_planes<N>, dim3((unsigned)((F + 255) / 256)), dim3(256), 0, s, F, m.fac.as<int32_t>(),
                       vu.as<double>(), sp.planes.as<float>());
    NSP_HIP(hipGetLastError());
    const int64_t ntiles = (cells + SR_SCAN_TILE - 1) / SR_SCAN_TILE;
    sp.count.alloc(sizeof(uint32_t) * cells);
    sp.start.alloc(sizeof(uint32_t) * (cells + 1));
    sp.tiles.alloc(sizeof(uint32_t) * (ntiles + 1));
    NSP_HIP(hipMemsetAsync(sp.count.p, 0, sizeof(uint32_t) * cells, s));
    NSP_HIP(hipStreamSynchronize(s));
    sp.n = N;
    sp.F = F;
    for (int k = 0; k < N; ++k) sp.light[k] = light[k];
    sp.ok = true;
}

template <int N>
void launch_t(const ShadowRasterPrep& sp, const TreeArgs& A, ShadowQueue<float> sq, int64_t scap,
              const unsigned long long* dc, RenderBuffers& b, float* Ls, int64_t cap, float* direct,
              const float* light32, float offset, unsigned long long* rays, hipStream_t s)
{
    SRLight lt{};
    for (int k = 0; k < N; ++k) lt.pos[k] = light32[k];
    lt.offset = offset;
    const SRGrid& g = sp.grid;
    const unsigned gq = 4096;  // a multiple of QSEG
    uint32_t* count = sp.count.as<uint32_t>();
    hipLaunchKernelGGL(k_sr_count<N>, dim3(gq), dim3(SR_BLOCK), 0, s, sq, scap, dc, g, b.sr_cell.as<int32_t>(),
                       b.sr_rank.as<uint32_t>(), count);
    NSP_HIP(hipGetLastError());
    const int64_t ntiles = (sp.cells + SR_SCAN_TILE - 1) / SR_SCAN_TILE;
    uint32_t* tiles = sp.tiles.as<uint32_t>();
    hipLaunchKernelGGL(k_sr_tile_sums, dim3((unsigned)ntiles), dim3(256), 0, s, count, sp.cells, tiles);
    NSP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sr_tile_scan, dim3(1), dim3(1024), 0, s, tiles, (int)ntiles, tiles + ntiles);
    NSP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sr_tile_down, dim3((unsigned)ntiles), dim3(256), 0, s, count, sp.cells, tiles,
                       tiles + ntiles, sp.start.as<uint32_t>());
    NSP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sr_scatter<N>, dim3(gq), dim3(SR_BLOCK), 0, s, sq, scap, dc, b.sr_cell.as<int32_t>(),
                       b.sr_rank.as<uint32_t>(), sp.start.as<uint32_t>(), b.sr_rays.as<float>(),
                       b.sr_ray.as<int32_t>(), b.sr_occq.as<uint8_t>());
    NSP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sr_raster<N>, dim3(NSP_SR_RASTER_BLOCKS), dim3(SR_BLOCK), 0, s, A, sp.unit.as<int2>(), sp.units,
                       sp.box.as<int32_t>(), sp.planes.as<float>(), g, sp.start.as<uint32_t>(),
                       b.sr_rays.as<float>(), scap, b.sr_ray.as<int32_t>(),
                       b.sr_occq.as<uint8_t>(), lt);
    NSP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sr_finish<N>, dim3(gq), dim3(SR_BLOCK), 0, s, sq, scap, dc, sp.cells,
                       sp.start.as<uint32_t>(), b.sr_occq.as<uint8_t>(), Ls, cap, direct, rays);
    NSP_HIP(hipGetLastError());
}

}  // namespace

bool shadow_raster_prepare(ShadowRasterPrep& sp, const MeshDev& m, const double* light, hipStream_t s)
{
    bool same = sp.n == m.n && sp.F == m.F && sp.planes.p;
    for (int k = 0; k < m.n && same; ++k) same = sp.light[k] == light[k];
    if (same) return sp.ok;
    switch (m.n) {
    case 3: prepare_t<3>(sp, m, light, s); break;
    case 4: prepare_t<4>(sp, m, light, s); break;
    case 5: prepare_t<5>(sp, m, light, s); break;
    case 6: prepare_t<6>(sp, m, light, s); break;
    default: sp.ok = false;
    }
    if (!sp.ok) {
        sp.n = m.n;
        sp.F = m.F;
        for (int k = 0; k < m.n; ++k) sp.light[k] = light[k];
        sp.planes.alloc(16);
    }
    return sp.ok;
}

void shadow_raster_buffers(RenderBuffers& b, int n, int64_t scap)
{
    b.sr_cell.alloc(sizeof(int32_t) * scap);
    b.sr_rank.alloc(sizeof(uint32_t) * scap);
    b.sr_rays.alloc(sizeof(float) * sr_record(n) * scap);
    b.sr_ray.alloc(sizeof(int32_t) * scap);
    b.sr_occq.alloc(scap);
}

void shadow_raster_launch(const ShadowRasterPrep& sp, const MeshDev& m, const TreeArgs& A, ShadowQueue<float> sq,
                          int64_t scap, const unsigned long long* dc, RenderBuffers& b, float* Ls, int64_t cap,
                          float* direct, const float* light32, float offset, unsigned long long* rays,
                          hipStream_t s)
{
    switch (m.n) {
    case 3: launch_t<3>(sp, A, sq, scap, dc, b, Ls, cap, direct, light32, offset, rays, s); break;
    case 4: launch_t<4>(sp, A, sq, scap, dc, b, Ls, cap, direct, light32, offset, rays, s); break;
    case 5: launch_t<5>(sp, A, sq, scap, dc, b, Ls, cap, direct, light32, offset, rays, s); break;
    case 6: launch_t<6>(sp, A, sq, scap, dc, b, Ls, cap, direct, light32, offset, rays, s); break;
    default: raise(NSP_E_INVALID, "dimension not supported");
    }
}

}  // namespace nsp
