// nsp_denoise.hip -- learned denoiser for 3D screens (n = 4 renders): an
// 8-layer 3x3x3 convolutional network, 32 channels, BF16 weights and
// activations, FP32 accumulation, trained on the GPU with Adam.
//
// No reference counterpart: the reference renders progressively and never
// denoises.  Every convolution is an implicit GEMM on v_mfma_f32_16x16x32_bf16
// (gfx950):
//   forward        Y[v, co]  = sum_{tap, ci} W[co, tap, ci] X[v + off(tap), ci]
//                  M = voxels (64 per wave), N = output channels, K = 27 * Cin
//   backward data  dX[v, ci] = sum_{tap, co} W[co, 26 - tap, ci] dY'[v + off(tap), co]
//                  the forward kernel on tap-flipped, transposed weights, with
//                  the ReLU mask of the previous layer applied in the epilogue
//   backward wgt   dW[co, tap, ci] = sum_v dY'[v, co] X[v + off(tap), ci]
//                  M = co, N = ci, K = voxels; per-chunk partials reduced in a
//                  fixed order, so gradients are bit-reproducible.
// Activations are channels-last BF16 ([voxel][channel]), so a lane's K slice
// of 8 channels is one 16-byte load and 16 lanes read 16 consecutive voxels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "nsp.h"
#include "nsp_device.hpp"
#include "nsp_internal.hpp"
#include "nsp_render.hpp"

namespace nsp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short shortx8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int DN_LAYERS = 8;
constexpr int DN_WIDTH = 32;
constexpr int DN_IN = 16;   // input channels, padded
constexpr int DN_OUT = 16;  // last layer output channels of the residual head, padded (3 used)
constexpr int DN_KOUT = 32; // last layer output channels of the kernel head, padded (27 used)

struct Dims {
    int nb, d, h, w;
    __host__ __device__ int64_t voxels() const { return (int64_t)nb * d * h * w; }
};

__device__ __forceinline__ bf16x8 zero8()
{
    shortx8 z = {0, 0, 0, 0, 0, 0, 0, 0};
    return __builtin_bit_cast(bf16x8, z);
}

__device__ __forceinline__ unsigned short f2bf(float f)
{
    __bf16 b = (__bf16)f;  // round to nearest even
    return __builtin_bit_cast(unsigned short, b);
}

__device__ __forceinline__ float bf2f(unsigned short u)
{
    return __builtin_bit_cast(float, (unsigned)u << 16);
}

// voxel index <-> coordinates (x fastest)
struct Vox {
    int b, z, y, x;
};
__device__ __forceinline__ Vox decode(int64_t v, const Dims& D)
{
    Vox r;
    r.x = (int)(v % D.w);
    v /= D.w;
    r.y = (int)(v % D.h);
    v /= D.h;
    r.z = (int)(v % D.d);
    r.b = (int)(v / D.d);
    return r;
}
__device__ __forceinline__ int64_t neighbour(const Vox& p, int tap, const Dims& D)
{
    int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    int z = p.z + dz, y = p.y + dy, x = p.x + dx;
    if (z < 0 || z >= D.d || y < 0 || y >= D.h || x < 0 || x >= D.w) return -1;
    return (((int64_t)p.b * D.d + z) * D.h + y) * D.w + x;
}

// ---------------------------------------------------------------------------
// Implicit-GEMM convolution (forward, and backward-data with flipped weights).
// A block computes a 16 (x) x 8 (y) x 4 (z) tile of output voxels.  Its input
// halo, 18 x 10 x 6 voxels x CIN channels, is staged once in LDS (8-channel
// groups outermost) and read there for all 27 taps.  Wave w takes the z-slice w: eight M-tiles (the x-rows
// y = 0..7, a lane row per voxel) by all output channels (NT N-tiles of 16).
// The weights are packed in MFMA fragment order, [K-step][N-tile][lane][8],
// so each lane's B fragment is one 16-byte load (the 54 KB of one layer stay
// in L2), prefetched a K-step ahead.  The epilogue stages the tile through LDS
// and stores 8 channels of a voxel per 16-byte store.
//   X      [V][CIN] bf16
//   Wf     fragment-ordered weights, KS = ceil(27 CIN / 32) K-steps of 32
//   bias   [NT*16] f32 or null
//   mask   [V][NT*16] bf16 or null: zero the output where mask <= 0 (ReLU' of the
//          layer whose gradient this is)
//   out    bf16 [V][NT*16] (out_f32 null) or f32 [V][NT*16]
constexpr int CT_X = 16, CT_Y = 8, CT_Z = 4;
constexpr int CH_X = CT_X + 2, CH_Y = CT_Y + 2, CH_Z = CT_Z + 2;
constexpr int CT_VOX = CT_X * CT_Y * CT_Z;
constexpr int CH_VOX = CH_X * CH_Y * CH_Z;

// tiles per volume along x, y, z
struct TileGrid {
    int tx, ty, tz;
};

// the epilogue's voxel stride in floats: COUT + 4 puts the four row groups of a
// wave's writes in different banks and keeps 16-byte alignment
template <int NT>
constexpr int epi_stride()
{
    return NT * 16 + 4;
}

template <int CIN, int NT>
constexpr int conv_lds_bytes()
{
    return CH_VOX * CIN * 2 > CT_VOX * epi_stride<NT>() * 4 ? CH_VOX * CIN * 2 : CT_VOX * epi_stride<NT>() * 4;
}

// Stages n_vox voxels of CIN channels (8-channel groups outermost) from global
// memory into LDS.  Unit u: a 16-lane group takes 16 consecutive voxels of one
// group, so a wave reads 16 voxels x 64 bytes contiguously and each group
// writes 256 consecutive bytes of LDS; every load of a thread is issued before
// its stores.  src(hv) gives the voxel's global index, or -1 outside the volume.
template <int CIN, int NVOX, int THREADS, typename Src>
__device__ __forceinline__ void stage_groups(const unsigned short* __restrict__ X, unsigned short* lds, Src src)
{
    constexpr int UPV = CIN / 8;
    constexpr int NU = ((NVOX + 15) / 16) * 16 * UPV;
    constexpr int ITER = (NU + THREADS - 1) / THREADS;
    uint4 buf[ITER];
#pragma unroll
    for (int k = 0; k < ITER; ++k) {
        const int u = (int)threadIdx.x + k * THREADS;
        const int hv = (u >> 4) / UPV * 16 + (u & 15), part = (u >> 4) % UPV;
        buf[k] = uint4{0u, 0u, 0u, 0u};
        if (u < NU && hv < NVOX) {
            const int64_t g = src(hv);
            if (g >= 0) buf[k] = *reinterpret_cast<const uint4*>(X + g * CIN + part * 8);
        }
    }
#pragma unroll
    for (int k = 0; k < ITER; ++k) {
        const int u = (int)threadIdx.x + k * THREADS;
        const int hv = (u >> 4) / UPV * 16 + (u & 15), part = (u >> 4) % UPV;
        if (u < NU && hv < NVOX) *reinterpret_cast<uint4*>(lds + (part * NVOX + hv) * 8) = buf[k];
    }
}

template <int CIN, int NT>
__global__ void __launch_bounds__(256) k_conv(const unsigned short* __restrict__ X, Dims D, TileGrid G,
                                              const unsigned short* __restrict__ Wf,
                                              const float* __restrict__ bias, int relu,
                                              const unsigned short* __restrict__ mask,
                                              unsigned short* __restrict__ out, float* __restrict__ out_f32)
{
    constexpr int KS = (27 * CIN + 31) / 32;
    constexpr int COUT = NT * 16;
    constexpr int ES = epi_stride<NT>();
    __shared__ __align__(16) unsigned char lds[conv_lds_bytes<CIN, NT>()];
    unsigned short* halo = reinterpret_cast<unsigned short*>(lds);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int t = blockIdx.x;
    const int tx = t % G.tx;
    t /= G.tx;
    const int ty = t % G.ty;
    t /= G.ty;
    const int tz = t % G.tz;
    const int b = t / G.tz;
    const int x0 = tx * CT_X, y0 = ty * CT_Y, z0 = tz * CT_Z;
    // channel-group-major halo: the 16 lanes of an A read take 16 consecutive
    // voxels of one 8-channel group, 16 consecutive bytes apart
    stage_groups<CIN, CH_VOX, 256>(X, halo, [&](int hv) -> int64_t {
        const int hx = hv % CH_X, hy = (hv / CH_X) % CH_Y, hz = hv / (CH_X * CH_Y);
        const int x = x0 + hx - 1, y = y0 + hy - 1, z = z0 + hz - 1;
        if (x < 0 || x >= D.w || y < 0 || y >= D.h || z < 0 || z >= D.d) return -1;
        return (((int64_t)b * D.d + z) * D.h + y) * D.w + x;
    });
    __syncthreads();
    const int row = lane & 15, kg = lane >> 4;
    floatx4 acc[CT_Y][NT];
#pragma unroll
    for (int m = 0; m < CT_Y; ++m)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[m][nt] = floatx4{0, 0, 0, 0};
    const bf16x8* W8 = reinterpret_cast<const bf16x8*>(Wf);
    bf16x8 bn[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) bn[nt] = W8[nt * 64 + lane];
    for (int s = 0; s < KS; ++s) {
        bf16x8 bc[NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) bc[nt] = bn[nt];
        if (s + 1 < KS)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) bn[nt] = W8[((s + 1) * NT + nt) * 64 + lane];
        const int k0 = s * 32 + kg * 8;
        const int tap_raw = k0 / CIN, c0 = k0 % CIN;
        // the padded end of K (CIN = 16: lanes 32-63 of the last step) reads tap 26
        // in bounds and multiplies zero
        const bool real = tap_raw < 27;
        const int tap = real ? tap_raw : 26;
        const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
        const unsigned short* src = halo + ((c0 / 8) * CH_VOX + ((wave + dz) * CH_Y + dy) * CH_X + row + dx) * 8;
#pragma unroll
        for (int m = 0; m < CT_Y; ++m) {
            bf16x8 a = *reinterpret_cast<const bf16x8*>(src + m * CH_X * 8);
            if (!real) a = zero8();
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                acc[m][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bc[nt], acc[m][nt], 0, 0, 0);
        }
    }
    __syncthreads();  // the halo is read; its LDS now holds the output tile
    float* epi = reinterpret_cast<float*>(lds);
    // D layout: lane holds rows 4*kg + i (x) of column `row` (channel) of each tile
#pragma unroll
    for (int m = 0; m < CT_Y; ++m)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                epi[((wave * CT_Y + m) * CT_X + kg * 4 + i) * ES + nt * 16 + row] = acc[m][nt][i];
    __syncthreads();
    constexpr int UPO = COUT / 8;
    for (int u = tid; u < CT_VOX * UPO; u += 256) {
        const int vl = u / UPO, part = u % UPO;
        const int x = x0 + vl % CT_X, y = y0 + (vl / CT_X) % CT_Y, z = z0 + vl / (CT_X * CT_Y);
        if (x >= D.w || y >= D.h || z >= D.d) continue;
        const int64_t v = (((int64_t)b * D.d + z) * D.h + y) * D.w + x;
        const float4 lo = *reinterpret_cast<const float4*>(epi + vl * ES + part * 8);
        const float4 hi = *reinterpret_cast<const float4*>(epi + vl * ES + part * 8 + 4);
        float r[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
        if (bias)
#pragma unroll
            for (int j = 0; j < 8; ++j) r[j] += bias[part * 8 + j];
        if (relu)
#pragma unroll
            for (int j = 0; j < 8; ++j) r[j] = fmaxf(r[j], 0.0f);
        if (mask) {
            const uint4 mk = *reinterpret_cast<const uint4*>(mask + v * COUT + part * 8);
            const unsigned mw[4] = {mk.x, mk.y, mk.z, mk.w};
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (!(bf2f((unsigned short)(mw[j / 2] >> (16 * (j & 1)))) > 0.0f)) r[j] = 0.0f;
        }
        if (out_f32) {
            *reinterpret_cast<float4*>(out_f32 + v * COUT + part * 8) = float4{r[0], r[1], r[2], r[3]};
            *reinterpret_cast<float4*>(out_f32 + v * COUT + part * 8 + 4) = float4{r[4], r[5], r[6], r[7]};
        } else {
            uint4 o;
            o.x = f2bf(r[0]) | ((unsigned)f2bf(r[1]) << 16);
            o.y = f2bf(r[2]) | ((unsigned)f2bf(r[3]) << 16);
            o.z = f2bf(r[4]) | ((unsigned)f2bf(r[5]) << 16);
            o.w = f2bf(r[6]) | ((unsigned)f2bf(r[7]) << 16);
            *reinterpret_cast<uint4*>(out + v * COUT + part * 8) = o;
        }
    }
}

// ---------------------------------------------------------------------------
// Weight and bias gradients: dW[co][tap][ci] = sum_v dY[v, co] X[v + off(tap), ci],
// db[co] = sum_v dY[v, co].  K runs over voxels, so both operands are staged in
// LDS transposed, voxels contiguous: dY of a 16 x 8 x 2 tile as [co][voxel], and
// X of its halo as [ci][z][y][x] with x rows padded to 24.  Wave w takes the
// nine taps with dz = w; for each (dy, ci tile) two aligned 16-byte reads hold
// x0..x0+15 of the shifted row, which give the B fragments of dx = 0, 1 and 2
// (the latter two by 2- and 4-byte shifts).  The bias gradient is one more
// MFMA against a B of ones (wave 0).  Blocks loop over tiles (tile = block +
// k * grid) and write one partial per block; k_reduce_dw sums the partials in
// a fixed order, so the gradients are bit-reproducible.
constexpr int WT_X = 16, WT_Y = 8, WT_Z = 2;
constexpr int WH_X = 24, WH_Y = WT_Y + 2, WH_Z = WT_Z + 2;  // the halo's x rows: 18 used, padded to 24
constexpr int WT_VOX = WT_X * WT_Y * WT_Z;
constexpr int WH_PLANE = WH_X * WH_Y * WH_Z;
// channel rows padded by 16 bytes: the 16 lanes of a fragment read 16 channel
// rows, which then start in different LDS banks
constexpr int WT_ROW = WT_VOX + 8;
constexpr int WH_ROW = WH_PLANE + 8;
constexpr int DW_MAX_BLOCKS = 512;

// Stages n_vox voxels of C channels from global memory into LDS transposed,
// [channel][voxel] with row stride ROW, for a block of 192 threads: a 16-lane
// group loads 16 consecutive voxels of one 8-channel group, and loads are
// issued in batches of 5 before their stores.  src(v): global voxel index or -1.
template <int C, int NVOX, int ROW, typename Src>
__device__ __forceinline__ void stage_transposed(const unsigned short* __restrict__ X, unsigned short* lds, Src src)
{
    constexpr int UPV = C / 8;
    constexpr int NU = ((NVOX + 15) / 16) * 16 * UPV;
    constexpr int BATCH = 5;
#pragma unroll 1
    for (int u0 = 0; u0 < NU; u0 += 192 * BATCH) {
        uint4 buf[BATCH];
#pragma unroll
        for (int k = 0; k < BATCH; ++k) {
            const int u = u0 + (int)threadIdx.x + k * 192;
            const int hv = (u >> 4) / UPV * 16 + (u & 15), part = (u >> 4) % UPV;
            buf[k] = uint4{0u, 0u, 0u, 0u};
            if (u < NU && hv < NVOX) {
                const int64_t g = src(hv);
                if (g >= 0) buf[k] = *reinterpret_cast<const uint4*>(X + g * C + part * 8);
            }
        }
#pragma unroll
        for (int k = 0; k < BATCH; ++k) {
            const int u = u0 + (int)threadIdx.x + k * 192;
            const int hv = (u >> 4) / UPV * 16 + (u & 15), part = (u >> 4) % UPV;
            if (u < NU && hv < NVOX) {
                const unsigned w4[4] = {buf[k].x, buf[k].y, buf[k].z, buf[k].w};
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    lds[(part * 8 + j) * ROW + hv] = (unsigned short)(w4[j / 2] >> (16 * (j & 1)));
            }
        }
    }
}

template <int CIN, int NT>
__global__ void __launch_bounds__(192) k_conv_dw(const unsigned short* __restrict__ X, Dims D, TileGrid G, int tiles,
                                                 const unsigned short* __restrict__ dY, float* __restrict__ partial,
                                                 float* __restrict__ bias_partial)
{
    constexpr int COUT = NT * 16;
    constexpr int CT = CIN / 16;
    __shared__ __align__(16) unsigned short xT[CIN * WH_ROW];
    __shared__ __align__(16) unsigned short dyT[COUT * WT_ROW];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;  // wave = dz
    const int row = lane & 15, kg = lane >> 4;
    floatx4 acc[9][NT][CT];
    floatx4 accb[NT];
#pragma unroll
    for (int q = 0; q < 9; ++q)
#pragma unroll
        for (int a = 0; a < NT; ++a)
#pragma unroll
            for (int c = 0; c < CT; ++c) acc[q][a][c] = floatx4{0, 0, 0, 0};
#pragma unroll
    for (int a = 0; a < NT; ++a) accb[a] = floatx4{0, 0, 0, 0};
    shortx8 ones8;
#pragma unroll
    for (int j = 0; j < 8; ++j) ones8[j] = (short)0x3F80;  // bf16 1.0
    const bf16x8 ones = __builtin_bit_cast(bf16x8, ones8);
    for (int t = blockIdx.x; t < tiles; t += gridDim.x) {
        int q = t;
        const int tx = q % G.tx;
        q /= G.tx;
        const int ty = q % G.ty;
        q /= G.ty;
        const int tz = q % G.tz;
        const int b = q / G.tz;
        const int x0 = tx * WT_X, y0 = ty * WT_Y, z0 = tz * WT_Z;
        stage_transposed<CIN, WH_PLANE, WH_ROW>(X, xT, [&](int hv) -> int64_t {
            const int hx = hv % WH_X, hy = (hv / WH_X) % WH_Y, hz = hv / (WH_X * WH_Y);
            const int x = x0 + hx - 1, y = y0 + hy - 1, z = z0 + hz - 1;
            if (hx >= WT_X + 2 || x < 0 || x >= D.w || y < 0 || y >= D.h || z < 0 || z >= D.d) return -1;
            return (((int64_t)b * D.d + z) * D.h + y) * D.w + x;
        });
        stage_transposed<COUT, WT_VOX, WT_ROW>(dY, dyT, [&](int vl) -> int64_t {
            const int x = x0 + vl % WT_X, y = y0 + (vl / WT_X) % WT_Y, z = z0 + vl / (WT_X * WT_Y);
            if (x >= D.w || y >= D.h || z >= D.d) return -1;
            return (((int64_t)b * D.d + z) * D.h + y) * D.w + x;
        });
        __syncthreads();
        for (int c = 0; c < WT_VOX / 32; ++c) {
            // this lane's 8 voxels: x-row r = z * WT_Y + y, x = xs .. xs + 7
            const int r = 2 * c + (kg >> 1), xs = (kg & 1) * 8;
            const int yy = r % WT_Y, zz = r / WT_Y;
            bf16x8 A[NT];
#pragma unroll
            for (int a = 0; a < NT; ++a)
                A[a] = *reinterpret_cast<const bf16x8*>(dyT + (a * 16 + row) * WT_ROW + r * WT_X + xs);
            if (wave == 0)
#pragma unroll
                for (int a = 0; a < NT; ++a)
                    accb[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[a], ones, accb[a], 0, 0, 0);
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int ct = 0; ct < CT; ++ct) {
                    const unsigned short* p =
                        xT + (ct * 16 + row) * WH_ROW + ((zz + wave) * WH_Y + yy + dy) * WH_X + xs;
                    const uint4 lo = *reinterpret_cast<const uint4*>(p);
                    const uint4 hi = *reinterpret_cast<const uint4*>(p + 8);
                    bf16x8 B[3];
                    B[0] = __builtin_bit_cast(bf16x8, lo);
                    uint4 s1, s2;
                    s1.x = (lo.x >> 16) | (lo.y << 16);
                    s1.y = (lo.y >> 16) | (lo.z << 16);
                    s1.z = (lo.z >> 16) | (lo.w << 16);
                    s1.w = (lo.w >> 16) | (hi.x << 16);
                    s2.x = lo.y;
                    s2.y = lo.z;
                    s2.z = lo.w;
                    s2.w = hi.x;
                    B[1] = __builtin_bit_cast(bf16x8, s1);
                    B[2] = __builtin_bit_cast(bf16x8, s2);
#pragma unroll
                    for (int dx = 0; dx < 3; ++dx)
#pragma unroll
                        for (int a = 0; a < NT; ++a)
                            acc[dy * 3 + dx][a][ct] =
                                __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[a], B[dx], acc[dy * 3 + dx][a][ct], 0, 0, 0);
                }
        }
        __syncthreads();
    }
    // D layout: lane holds rows 4*kg + i (co) of column `row` (ci)
#pragma unroll
    for (int q = 0; q < 9; ++q)
#pragma unroll
        for (int a = 0; a < NT; ++a)
#pragma unroll
            for (int c = 0; c < CT; ++c)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int co = a * 16 + kg * 4 + i, ci = c * 16 + row, tap = wave * 9 + q;
                    partial[(((int64_t)blockIdx.x * COUT + co) * 27 + tap) * CIN + ci] = acc[q][a][c][i];
                }
    if (wave == 0 && row == 0)
#pragma unroll
        for (int a = 0; a < NT; ++a)
#pragma unroll
            for (int i = 0; i < 4; ++i) bias_partial[(int64_t)blockIdx.x * COUT + a * 16 + kg * 4 + i] = accb[a][i];
}

// dW = the sum of the blocks' partials (also of the bias partials), in a fixed
// order: a block of 256 threads takes 64 elements; its four waves sum four
// fixed ranges of partials in order (loads coalesced over the elements), and
// the four range sums are added in range order.
__global__ void __launch_bounds__(256) k_reduce_dw(const float* __restrict__ partial, int blocks, int n,
                                                   float* __restrict__ dw)
{
    __shared__ float part[4][64];
    const int col = threadIdx.x & 63, r = threadIdx.x >> 6;
    const int e = blockIdx.x * 64 + col;
    const int per = (blocks + 3) / 4;
    const int c1 = min(blocks, (r + 1) * per);
    float s = 0;
    if (e < n)
        for (int c = r * per; c < c1; ++c) s += partial[(int64_t)c * n + e];
    part[r][col] = s;
    __syncthreads();
    if (r == 0 && e < n) dw[e] = ((part[0][col] + part[1][col]) + part[2][col]) + part[3][col];
}

// ---------------------------------------------------------------------------
// Loss (mean squared error of prediction = net + residual radiance) and its gradient.
__global__ void k_loss(const float* __restrict__ out, const float* __restrict__ resid, const float* __restrict__ tgt,
                       int64_t V, float scale, unsigned short* __restrict__ dout, float* __restrict__ pred,
                       float* __restrict__ loss_part)
{
    __shared__ float sm[256];
    float l = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        for (int c = 0; c < DN_OUT; ++c) {
            if (c < 3) {
                float p = out[v * DN_OUT + c] + (resid ? resid[v * 3 + c] : 0.0f);
                if (pred) pred[v * 3 + c] = p;
                if (tgt) {
                    float e = p - tgt[v * 3 + c];
                    l += e * e;
                    if (dout) dout[v * DN_OUT + c] = f2bf(2.0f * e * scale);
                }
            } else if (dout) {
                dout[v * DN_OUT + c] = 0;
            }
        }
    }
    sm[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0 && loss_part) loss_part[blockIdx.x] = sm[0];
}

// Kernel head: out[v][0..27) are the logits of a 3x3x3 filter; the prediction
// is the softmax-weighted sum (over the taps inside the volume) of the 3
// radiance channels rad[.][3] of the tap voxels.  With tgt: squared error
// into loss_part and, with dout, the logits' gradient (bf16 [V][32]):
// dL/dz_t = w_t (g_t - sum_s w_s g_s), g_t = dL/dpred . rad(tap t).
__global__ void k_kp_loss(const float* __restrict__ out, const float* __restrict__ rad,
                          const float* __restrict__ tgt, Dims D, float scale, unsigned short* __restrict__ dout,
                          float* __restrict__ pred, float* __restrict__ loss_part)
{
    __shared__ float sm[256];
    const int64_t V = D.voxels();
    float l = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        const Vox p = decode(v, D);
        float z[27];
        int64_t nb[27];
        float m = -INFINITY;
#pragma unroll
        for (int t = 0; t < 27; ++t) {
            nb[t] = neighbour(p, t, D);
            z[t] = out[v * DN_KOUT + t];
            if (nb[t] >= 0) m = fmaxf(m, z[t]);
        }
        float sum = 0;
#pragma unroll
        for (int t = 0; t < 27; ++t) {
            z[t] = nb[t] >= 0 ? expf(z[t] - m) : 0.0f;
            sum += z[t];
        }
        const float inv = 1.0f / sum;
        float pr[3] = {0, 0, 0};
#pragma unroll
        for (int t = 0; t < 27; ++t) {
            z[t] *= inv;  // weight
            if (nb[t] >= 0)
#pragma unroll
                for (int c = 0; c < 3; ++c) pr[c] += z[t] * rad[nb[t] * 3 + c];
        }
        if (pred)
#pragma unroll
            for (int c = 0; c < 3; ++c) pred[v * 3 + c] = pr[c];
        if (tgt) {
            float gp[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float e = pr[c] - tgt[v * 3 + c];
                l += e * e;
                gp[c] = 2.0f * e * scale;
            }
            if (dout) {
                float g[27], wg = 0;
#pragma unroll
                for (int t = 0; t < 27; ++t) {
                    g[t] = 0;
                    if (nb[t] >= 0)
#pragma unroll
                        for (int c = 0; c < 3; ++c) g[t] += gp[c] * rad[nb[t] * 3 + c];
                    wg += z[t] * g[t];
                }
#pragma unroll
                for (int t = 0; t < 27; ++t) dout[v * DN_KOUT + t] = f2bf(nb[t] >= 0 ? z[t] * (g[t] - wg) : 0.0f);
#pragma unroll
                for (int t = 27; t < DN_KOUT; ++t) dout[v * DN_KOUT + t] = 0;
            }
        }
    }
    sm[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0 && loss_part) loss_part[blockIdx.x] = sm[0];
}

constexpr int LOSS_BLOCKS = 1024;

// One step's loss: the blocks' partial sums in a fixed order (each of 256
// threads sums the partials b = thread (mod 256) in order, then a tree), times scale.
__global__ void __launch_bounds__(256) k_loss_total(const float* __restrict__ part, int blocks, float scale,
                                                    float* __restrict__ out)
{
    __shared__ double sm[256];
    double s = 0;
    for (int b = threadIdx.x; b < blocks; b += 256) s += part[b];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sm[threadIdx.x] += sm[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = (float)(sm[0] * scale);
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       int64_t n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    float mi = b1 * m[i] + (1 - b1) * gi;
    float vi = b2 * v[i] + (1 - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Pack layer weights W[co][tap][ci] (FP32 master, padded shapes) into the BF16
// fragment order of k_conv, [K-step][N-tile][lane][8]: for the forward
// convolution element (s, nt, l, j) is W[co][tap][ci] with co = 16 nt + (l & 15),
// k = 32 s + 8 (l >> 4) + j = tap * cin + ci (zero for tap >= 27); for the
// backward-data convolution (cin' = cout, cout' = cin, taps flipped) it is
// W[c][26 - tap'][o] with o = 16 nt + (l & 15), k = tap' * cout + c.
__global__ void k_pack_w(const float* __restrict__ W, int cin, int cout, int kpf, int kpb,
                         unsigned short* __restrict__ wf, unsigned short* __restrict__ wb)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = e & 7, l = (e >> 3) & 63, q = e >> 9;
    if (e < cout * kpf) {
        const int nts = cout / 16, nt = q % nts, s = q / nts;
        const int co = nt * 16 + (l & 15), k = s * 32 + (l >> 4) * 8 + j;
        const int tap = k / cin, ci = k % cin;
        wf[e] = tap < 27 ? f2bf(W[((int64_t)co * 27 + tap) * cin + ci]) : (unsigned short)0;
    }
    if (wb && e < cin * kpb) {
        const int nts = cin / 16, nt = q % nts, s = q / nts;
        const int o = nt * 16 + (l & 15), k = s * 32 + (l >> 4) * 8 + j;
        const int tp = k / cout, c = k % cout;
        wb[e] = tp < 27 ? f2bf(W[((int64_t)c * 27 + (26 - tp)) * cin + o]) : (unsigned short)0;
    }
}

// Input conversion: f32 [V][cin_real] -> bf16 [V][16]; optional residual copy.
__global__ void k_pack_input(const float* __restrict__ in, int64_t V, int cin, int resid_ch,
                             unsigned short* __restrict__ x, float* __restrict__ resid)
{
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    for (int c = 0; c < DN_IN; ++c) x[v * DN_IN + c] = c < cin ? f2bf(in[v * cin + c]) : 0;
    if (resid)
        for (int c = 0; c < 3; ++c) resid[v * 3 + c] = resid_ch >= 0 ? in[v * cin + resid_ch + c] : 0.0f;
}

// The denoiser input of a 3D-screen scene from its per-pixel sums in device
// memory (pixel order = voxel order, x fastest): channels [1-spp radiance (the
// first-sample sums over the passes), 4-spp radiance, albedo, normal (4)],
// each the mean over its samples, BF16, padded to 16; the residual is the
// 4-spp radiance.  aovc: offset of the first-sample channels in aov.
__global__ void k_scene_input(const float* __restrict__ sum, const int32_t* __restrict__ cnt,
                              const float* __restrict__ aov, int aovc, int64_t P, int spp_pass,
                              unsigned short* __restrict__ x, float* __restrict__ resid)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= P) return;
    const int c = cnt[v];
    const float inv = c > 0 ? 1.0f / (float)c : 0.0f;
    const int passes = spp_pass > 0 ? c / spp_pass : 0;
    const float inv1 = passes > 0 ? 1.0f / (float)passes : 0.0f;
    float in[16];
    for (int k = 0; k < 3; ++k) {
        in[k] = aov[(int64_t)(aovc + k) * P + v] * inv1;
        in[3 + k] = sum[(int64_t)k * P + v] * inv;
        in[6 + k] = aov[(int64_t)k * P + v] * inv;
        resid[v * 3 + k] = in[3 + k];
    }
    for (int k = 0; k < 4; ++k) in[9 + k] = aov[(int64_t)(3 + k) * P + v] * inv;
    for (int k = 13; k < 16; ++k) in[k] = 0.0f;
    for (int k = 0; k < 16; ++k) x[v * DN_IN + k] = f2bf(in[k]);
}

// pred = blend * pred + (1 - blend) * resid (the network's output blended with its input)
__global__ void k_blend(float* __restrict__ pred, const float* __restrict__ resid, int64_t n, float blend)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) pred[i] = blend * pred[i] + (1.0f - blend) * resid[i];
}

// Random training patches: for patch i of the batch, a Philox-drawn volume and
// corner; gathers inputs (f32 -> packed bf16), residual and target.
__global__ void k_gather_patches(const float* __restrict__ vin, const float* __restrict__ vtgt, int S, int Dd, int Hh,
                                 int Ww, int cin, int resid_ch, int P, int batch, uint64_t seed, int step,
                                 unsigned short* __restrict__ x, float* __restrict__ resid, float* __restrict__ tgt)
{
    int64_t V = (int64_t)batch * P * P * P;
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    int px = (int)(v % P), py = (int)((v / P) % P), pz = (int)((v / P / P) % P), bi = (int)(v / P / P / P);
    uint32_t o[4];
    philox4x32((uint32_t)bi, (uint32_t)step, 0x5EEDu, 0, (uint32_t)seed, (uint32_t)(seed >> 32), o);
    int s = (int)(o[0] % (uint32_t)S);
    int z0 = (int)(o[1] % (uint32_t)(Dd - P + 1)), y0 = (int)(o[2] % (uint32_t)(Hh - P + 1)),
        x0 = (int)(o[3] % (uint32_t)(Ww - P + 1));
    int64_t src = (((int64_t)s * Dd + (z0 + pz)) * Hh + (y0 + py)) * Ww + (x0 + px);
    for (int c = 0; c < DN_IN; ++c) x[v * DN_IN + c] = c < cin ? f2bf(vin[src * cin + c]) : 0;
    for (int c = 0; c < 3; ++c) {
        resid[v * 3 + c] = resid_ch >= 0 ? vin[src * cin + resid_ch + c] : 0.0f;
        tgt[v * 3 + c] = vtgt[src * 3 + c];
    }
}

// ---------------------------------------------------------------------------

struct Layer {
    int cin, cout, relu;
    int64_t w_off, b_off;  // offsets into the master parameter vector
    int kpf, kpb;
    DevBuf wf, wb;         // packed bf16
};

static int kp_of(int c) { return ((27 * c + 31) / 32) * 32; }

}  // namespace nsp

using namespace nsp;

struct nsp_context;  // defined in nsp_api.cpp

struct nsp_denoiser {
    nsp_context* ctx;
    int device;
    hipStream_t stream;
    std::string* err;
    int cin_real;
    int resid_ch;  // residual head: the residual channels; kernel head: the radiance channels it filters
    int head = NSP_DENOISER_RESIDUAL;
    int64_t nparams = 0;
    std::vector<Layer> layers;
    DevBuf params, grads, adam_m, adam_v;
    int64_t adam_t = 0;
    // activations of the last forward: act[0] = input (bf16 [V][16]), act[l+1] = layer l output
    DevBuf act[DN_LAYERS + 1];
    DevBuf out_f32, resid, target, pred, dy_a, dy_b, partial, bias_part, loss_part;
    int64_t cap_v = 0;
    // device time of the last forward's eight convolutions (HIP events on the stream)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_forward_ms = 0;
    int64_t last_forward_voxels = 0;
    int64_t last_output_voxels = 0;  // voxels of the prediction the last forward left in pred
    ~nsp_denoiser()
    {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
    }
};

namespace nsp {
// from nsp_api.cpp
hipStream_t context_stream(nsp_context* ctx);
int context_device(nsp_context* ctx);
std::string* context_error(nsp_context* ctx);
}  // namespace nsp

namespace {

void dn_alloc(nsp_denoiser& d, int64_t V)
{
    if (V <= d.cap_v) return;
    d.act[0].alloc(sizeof(unsigned short) * V * DN_IN);
    for (int l = 0; l < DN_LAYERS; ++l) d.act[l + 1].alloc(sizeof(unsigned short) * V * d.layers[l].cout);
    d.out_f32.alloc(sizeof(float) * V * d.layers[DN_LAYERS - 1].cout);
    d.resid.alloc(sizeof(float) * V * 3);
    d.target.alloc(sizeof(float) * V * 3);
    d.pred.alloc(sizeof(float) * V * 3);
    d.dy_a.alloc(sizeof(unsigned short) * V * DN_WIDTH);
    d.dy_b.alloc(sizeof(unsigned short) * V * DN_WIDTH);
    d.partial.alloc(sizeof(float) * DW_MAX_BLOCKS * 27 * DN_WIDTH * DN_WIDTH);
    d.bias_part.alloc(sizeof(float) * DW_MAX_BLOCKS * DN_WIDTH);
    d.loss_part.alloc(sizeof(float) * LOSS_BLOCKS);
    d.cap_v = V;
}

void pack_weights(nsp_denoiser& d)
{
    hipStream_t s = d.stream;
    for (Layer& L : d.layers) {
        int n = std::max(L.cout * L.kpf, L.cin * L.kpb);
        hipLaunchKernelGGL(k_pack_w, dim3((n + 255) / 256), dim3(256), 0, s, d.params.as<float>() + L.w_off, L.cin,
                           L.cout, L.kpf, L.kpb, L.wf.as<unsigned short>(), L.wb.as<unsigned short>());
    }
    NSP_HIP(hipGetLastError());
}

static TileGrid tiles_of(Dims D, int tx, int ty, int tz)
{
    return TileGrid{(D.w + tx - 1) / tx, (D.h + ty - 1) / ty, (D.d + tz - 1) / tz};
}

template <int CIN, int NT>
void conv_launch(const unsigned short* x, Dims D, const unsigned short* w, const float* bias, int relu,
                 const unsigned short* mask, unsigned short* out, float* out_f32, hipStream_t s)
{
    const TileGrid G = tiles_of(D, CT_X, CT_Y, CT_Z);
    const int64_t blocks = (int64_t)D.nb * G.tx * G.ty * G.tz;
    if (blocks > 0x7fffffff) raise(NSP_E_INVALID, "volume too large for the convolution grid");
    hipLaunchKernelGGL((k_conv<CIN, NT>), dim3((unsigned)blocks), dim3(256), 0, s, x, D, G, w, bias, relu, mask, out,
                       out_f32);
    NSP_HIP(hipGetLastError());
}

void conv(int cin, int cout, const unsigned short* x, Dims D, const unsigned short* w, const float* bias, int relu,
          const unsigned short* mask, unsigned short* out, float* out_f32, hipStream_t s)
{
    if (cin == 16 && cout == 32) return conv_launch<16, 2>(x, D, w, bias, relu, mask, out, out_f32, s);
    if (cin == 32 && cout == 32) return conv_launch<32, 2>(x, D, w, bias, relu, mask, out, out_f32, s);
    if (cin == 32 && cout == 16) return conv_launch<32, 1>(x, D, w, bias, relu, mask, out, out_f32, s);
    if (cin == 16 && cout == 16) return conv_launch<16, 1>(x, D, w, bias, relu, mask, out, out_f32, s);
    raise(NSP_E_INVALID, "unsupported conv shape");
}

// weight and bias gradient partials of one layer; returns the number of blocks (partials)
int conv_dw(int cin, int cout, const unsigned short* x, Dims D, const unsigned short* dy, float* partial,
            float* bias_partial, hipStream_t s)
{
    const TileGrid G = tiles_of(D, WT_X, WT_Y, WT_Z);
    const int64_t tiles = (int64_t)D.nb * G.tx * G.ty * G.tz;
    if (tiles > 0x7fffffff) raise(NSP_E_INVALID, "volume too large for the gradient grid");
    const int blocks = (int)std::min<int64_t>(tiles, DW_MAX_BLOCKS);
    if (cin == 16 && cout == 32)
        hipLaunchKernelGGL((k_conv_dw<16, 2>), dim3(blocks), dim3(192), 0, s, x, D, G, (int)tiles, dy, partial,
                           bias_partial);
    else if (cin == 32 && cout == 32)
        hipLaunchKernelGGL((k_conv_dw<32, 2>), dim3(blocks), dim3(192), 0, s, x, D, G, (int)tiles, dy, partial,
                           bias_partial);
    else if (cin == 32 && cout == 16)
        hipLaunchKernelGGL((k_conv_dw<32, 1>), dim3(blocks), dim3(192), 0, s, x, D, G, (int)tiles, dy, partial,
                           bias_partial);
    else
        raise(NSP_E_INVALID, "unsupported conv shape");
    NSP_HIP(hipGetLastError());
    return blocks;
}

// forward pass over act[0] (already packed) -> out_f32, pred (if requested)
void forward(nsp_denoiser& d, Dims D)
{
    hipStream_t s = d.stream;
    for (int l = 0; l < DN_LAYERS; ++l) {
        Layer& L = d.layers[l];
        bool last = l == DN_LAYERS - 1;
        conv(L.cin, L.cout, d.act[l].as<unsigned short>(), D, L.wf.as<unsigned short>(),
             d.params.as<float>() + L.b_off, L.relu, nullptr, last ? nullptr : d.act[l + 1].as<unsigned short>(),
             last ? d.out_f32.as<float>() : nullptr, s);
    }
}

// backward from dout (bf16 [V][16]) -> grads (FP32, master layout)
void backward(nsp_denoiser& d, Dims D, const unsigned short* dout)
{
    hipStream_t s = d.stream;
    const unsigned short* dy = dout;
    for (int l = DN_LAYERS - 1; l >= 0; --l) {
        Layer& L = d.layers[l];
        // weight and bias gradients: per-block partials [block][co][tap][ci] (the
        // master layout) and [block][co], summed in block order
        const int blocks =
            conv_dw(L.cin, L.cout, d.act[l].as<unsigned short>(), D, dy, d.partial.as<float>(), d.bias_part.as<float>(), s);
        int n = 27 * L.cout * L.cin;
        hipLaunchKernelGGL(k_reduce_dw, dim3((n + 63) / 64), dim3(256), 0, s, d.partial.as<float>(), blocks, n,
                           d.grads.as<float>() + L.w_off);
        hipLaunchKernelGGL(k_reduce_dw, dim3((L.cout + 63) / 64), dim3(256), 0, s, d.bias_part.as<float>(), blocks,
                           L.cout, d.grads.as<float>() + L.b_off);
        NSP_HIP(hipGetLastError());
        if (l == 0) break;
        // data gradient for layer l-1, masked by its ReLU
        unsigned short* next = (dy == d.dy_a.as<unsigned short>()) ? d.dy_b.as<unsigned short>()
                                                                     : d.dy_a.as<unsigned short>();
        conv(L.cout, L.cin, dy, D, L.wb.as<unsigned short>(), nullptr, 0, d.act[l].as<unsigned short>(), next,
             nullptr, s);
        dy = next;
    }
}

// The loss kernels of one step write LOSS_BLOCKS partial sums; with loss_out
// the step's mean squared error is summed into loss_out on the device (no host
// wait), else it is returned after a wait.
float loss_and_grad(nsp_denoiser& d, Dims D, bool want_grad, float* loss_out = nullptr)
{
    hipStream_t s = d.stream;
    const int64_t V = D.voxels();
    const int blocks = LOSS_BLOCKS;
    float scale = 1.0f / (float)(3 * V);
    if (d.head == NSP_DENOISER_KERNEL)
        hipLaunchKernelGGL(k_kp_loss, dim3(blocks), dim3(256), 0, s, d.out_f32.as<float>(), d.resid.as<float>(),
                           d.target.as<float>(), D, scale, want_grad ? d.dy_a.as<unsigned short>() : nullptr,
                           d.pred.as<float>(), d.loss_part.as<float>());
    else
        hipLaunchKernelGGL(k_loss, dim3(blocks), dim3(256), 0, s, d.out_f32.as<float>(), d.resid.as<float>(),
                           d.target.as<float>(), V, scale, want_grad ? d.dy_a.as<unsigned short>() : nullptr,
                           d.pred.as<float>(), d.loss_part.as<float>());
    NSP_HIP(hipGetLastError());
    if (loss_out) {
        hipLaunchKernelGGL(k_loss_total, dim3(1), dim3(256), 0, s, d.loss_part.as<float>(), blocks, scale, loss_out);
        NSP_HIP(hipGetLastError());
        return 0.0f;
    }
    std::vector<float> lp(blocks);
    NSP_HIP(hipMemcpyAsync(lp.data(), d.loss_part.p, sizeof(float) * blocks, hipMemcpyDeviceToHost, s));
    NSP_HIP(hipStreamSynchronize(s));
    double sum = 0;
    for (float x : lp) sum += x;
    return (float)(sum * scale);
}

}  // namespace

static Dims dims_of(int nb, int dd, int hh, int ww) { return Dims{nb, dd, hh, ww}; }

extern "C" {

int nsp_denoiser_create(nsp_context* ctx, int in_channels, int residual_channel, uint64_t seed, nsp_denoiser** out)
{
    return nsp_denoiser_create_head(ctx, in_channels, residual_channel, NSP_DENOISER_RESIDUAL, seed, out);
}

int nsp_denoiser_create_head(nsp_context* ctx, int in_channels, int residual_channel, int head, uint64_t seed,
                             nsp_denoiser** out)
{
    if (!ctx || !out) return NSP_E_INVALID;
    *out = nullptr;
    std::string* err = context_error(ctx);
    try {
        NSP_HIP(hipSetDevice(context_device(ctx)));
        if (in_channels < 1 || in_channels > DN_IN) raise(NSP_E_INVALID, "denoiser input channels must be in [1, 16]");
        if (head != NSP_DENOISER_RESIDUAL && head != NSP_DENOISER_KERNEL) raise(NSP_E_INVALID, "unknown head");
        if (residual_channel + 3 > in_channels || (head == NSP_DENOISER_KERNEL && residual_channel < 0))
            raise(NSP_E_INVALID, "residual channels out of range");
        auto d = std::make_unique<nsp_denoiser>();
        d->ctx = ctx;
        d->device = context_device(ctx);
        d->stream = context_stream(ctx);
        d->err = err;
        d->cin_real = in_channels;
        d->resid_ch = residual_channel;
        d->head = head;
        int64_t off = 0;
        for (int l = 0; l < DN_LAYERS; ++l) {
            Layer L;
            L.cin = l == 0 ? DN_IN : DN_WIDTH;
            L.cout = l == DN_LAYERS - 1 ? (head == NSP_DENOISER_KERNEL ? DN_KOUT : DN_OUT) : DN_WIDTH;
            L.relu = l < DN_LAYERS - 1;
            L.w_off = off;
            off += (int64_t)L.cout * 27 * L.cin;
            L.b_off = off;
            off += L.cout;
            L.kpf = kp_of(L.cin);
            L.kpb = kp_of(L.cout);
            d->layers.push_back(std::move(L));
        }
        d->nparams = off;
        // He-uniform initialisation over the real fan-in; padded entries stay zero.  The
        // kernel head's last layer starts at zero: uniform weights, a 3x3x3 box filter.
        std::vector<float> p(off, 0.0f);
        uint64_t st = seed;
        auto next = [&]() {
            uint64_t z = (st += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            return (double)(z >> 11) * (1.0 / 9007199254740992.0);
        };
        for (int l = 0; l < DN_LAYERS; ++l) {
            const Layer& L = d->layers[l];
            int ci_real = l == 0 ? in_channels : DN_WIDTH;
            int co_real = l == DN_LAYERS - 1 ? (head == NSP_DENOISER_KERNEL ? 0 : 3) : DN_WIDTH;
            double bound = std::sqrt(6.0 / (27.0 * ci_real)) * (l == DN_LAYERS - 1 ? 0.1 : 1.0);
            for (int co = 0; co < co_real; ++co)
                for (int t = 0; t < 27; ++t)
                    for (int ci = 0; ci < ci_real; ++ci)
                        p[L.w_off + ((int64_t)co * 27 + t) * L.cin + ci] = (float)((2 * next() - 1) * bound);
        }
        d->params.alloc(sizeof(float) * off);
        d->grads.alloc(sizeof(float) * off);
        d->adam_m.alloc(sizeof(float) * off);
        d->adam_v.alloc(sizeof(float) * off);
        NSP_HIP(hipMemcpy(d->params.p, p.data(), sizeof(float) * off, hipMemcpyHostToDevice));
        NSP_HIP(hipMemset(d->adam_m.p, 0, sizeof(float) * off));
        NSP_HIP(hipMemset(d->adam_v.p, 0, sizeof(float) * off));
        for (Layer& L : d->layers) {
            L.wf.alloc(sizeof(unsigned short) * L.cout * L.kpf);
            L.wb.alloc(sizeof(unsigned short) * L.cin * L.kpb);
        }
        pack_weights(*d);
        NSP_HIP(hipStreamSynchronize(d->stream));
        *out = d.release();
        return NSP_OK;
    } catch (const Error& e) {
        *err = e.msg;
        return e.code;
    }
}

void nsp_denoiser_destroy(nsp_denoiser* d)
{
    if (!d) return;
    (void)hipSetDevice(d->device);
    (void)hipDeviceSynchronize();
    delete d;
}

int64_t nsp_denoiser_param_count(const nsp_denoiser* d) { return d ? d->nparams : NSP_E_INVALID; }

int nsp_denoiser_get_params(const nsp_denoiser* d, float* out)
{
    if (!d || !out) return NSP_E_INVALID;
    try {
        NSP_HIP(hipSetDevice(d->device));
        NSP_HIP(hipMemcpy(out, d->params.p, sizeof(float) * d->nparams, hipMemcpyDeviceToHost));
        return NSP_OK;
    } catch (const Error& e) {
        *d->err = e.msg;
        return e.code;
    }
}

int nsp_denoiser_set_params(nsp_denoiser* d, const float* in)
{
    if (!d || !in) return NSP_E_INVALID;
    try {
        NSP_HIP(hipSetDevice(d->device));
        NSP_HIP(hipMemcpy(d->params.p, in, sizeof(float) * d->nparams, hipMemcpyHostToDevice));
        NSP_HIP(hipMemset(d->adam_m.p, 0, sizeof(float) * d->nparams));
        NSP_HIP(hipMemset(d->adam_v.p, 0, sizeof(float) * d->nparams));
        d->adam_t = 0;
        pack_weights(*d);
        NSP_HIP(hipStreamSynchronize(d->stream));
        return NSP_OK;
    } catch (const Error& e) {
        *d->err = e.msg;
        return e.code;
    }
}

int nsp_denoiser_get_state(const nsp_denoiser* d, float* params, float* adam_m, float* adam_v, int64_t* adam_t)
{
    if (!d) return NSP_E_INVALID;
    try {
        NSP_HIP(hipSetDevice(d->device));
        NSP_HIP(hipStreamSynchronize(d->stream));
        if (params) NSP_HIP(hipMemcpy(params, d->params.p, sizeof(float) * d->nparams, hipMemcpyDeviceToHost));
        if (adam_m) NSP_HIP(hipMemcpy(adam_m, d->adam_m.p, sizeof(float) * d->nparams, hipMemcpyDeviceToHost));
        if (adam_v) NSP_HIP(hipMemcpy(adam_v, d->adam_v.p, sizeof(float) * d->nparams, hipMemcpyDeviceToHost));
        if (adam_t) *adam_t = d->adam_t;
        return NSP_OK;
    } catch (const Error& e) {
        *d->err = e.msg;
        return e.code;
    }
}

int nsp_denoiser_set_state(nsp_denoiser* d, const float* params, const float* adam_m, const float* adam_v,
                           int64_t adam_t)
{
    if (!d || !params || !adam_m || !adam_v || adam_t < 0) return NSP_E_INVALID;
    try {
        NSP_HIP(hipSetDevice(d->device));
        NSP_HIP(hipStreamSynchronize(d->stream));
        NSP_HIP(hipMemcpy(d->params.p, params, sizeof(float) * d->nparams, hipMemcpyHostToDevice));
        NSP_HIP(hipMemcpy(d->adam_m.p, adam_m, sizeof(float) * d->nparams, hipMemcpyHostToDevice));
        NSP_HIP(hipMemcpy(d->adam_v.p, adam_v, sizeof(float) * d->nparams, hipMemcpyHostToDevice));
        d->adam_t = adam_t;
        pack_weights(*d);
        NSP_HIP(hipStreamSynchronize(d->stream));
        return NSP_OK;
    } catch (const Error& e) {
        *d->err = e.msg;
        return e.code;
    }
}

int nsp_denoiser_forward_scene(nsp_denoiser* d, const nsp_scene* scene, int samples_per_pass, float blend,
                               float* output)
{
    if (!d || !scene || samples_per_pass < 1 || !(blend >= 0.0f && blend <= 1.0f)) return NSP_E_INVALID;
    try {
        NSP_HIP(hipSetDevice(d->device));
        ScenePixels sp;
        if (!scene_pixel_buffers(scene, &sp)) raise(NSP_E_INVALID, "the scene has no AOV sums (paint it with aov = 1)");
        if (sp.n != 4) raise(NSP_E_UNSUPPORTED, "the denoiser takes the 3D screens of n = 4 scenes");
        if (d->cin_real != 13 || d->resid_ch != 3)
            raise(NSP_E_INVALID, "a scene's input has 13 channels (1-spp, 4-spp radiance, albedo, normal), residual 3");
        if (sp.ctx != d->ctx) raise(NSP_E_INVALID, "the scene belongs to another context");
        Dims D = dims_of(1, sp.screen[2], sp.screen[1], sp.screen[0]);
        const int64_t V = D.voxels();
        dn_alloc(*d, V);
        hipStream_t s = d->stream;
        hipLaunchKernelGGL(k_scene_input, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, s, sp.sum, sp.cnt, sp.aov,
                           aov_channels(4), V, samples_per_pass, d->act[0].as<unsigned short>(), d->resid.as<float>());
        NSP_HIP(hipGetLastError());
        if (!d->ev0) NSP_HIP(hipEventCreate(&d->ev0));
        if (!d->ev1) NSP_HIP(hipEventCreate(&d->ev1));
        NSP_HIP(hipEventRecord(d->ev0, s));
        forward(*d, D);
        NSP_HIP(hipEventRecord(d->ev1, s));
        if (d->head == NSP_DENOISER_KERNEL)
            hipLaunchKernelGGL(k_kp_loss, dim3(LOSS_BLOCKS), dim3(256), 0, s, d->out_f32.as<float>(),
                               d->resid.as<float>(), (const float*)nullptr, D, 1.0f, (unsigned short*)nullptr,
                               d->pred.as<float>(), (float*)nullptr);
        else
            hipLaunchKernelGGL(k_loss, dim3(LOSS_BLOCKS), dim3(256), 0, s, d->out_f32.as<float>(), d->resid.as<float>(),
                               (const float*)nullptr, V, 1.0f, (unsigned short*)nullptr, d->pred.as<float>(),
                               (float*)nullptr);
        if (blend < 1.0f)
            hipLaunchKernelGGL(k_blend, dim3((unsigned)((3 * V + 255) / 256)), dim3(256), 0, s, d->pred.as<float>(),
                               d->resid.as<float>(), 3 * V, blend);
        NSP_HIP(hipGetLastError());
        if (output) NSP_HIP(hipMemcpyAsync(output, d->pred.p, sizeof(float) * V * 3, hipMemcpyDeviceToHost, s));
        NSP_HIP(hipStreamSynchronize(s));
        float ms = 0;
        NSP_HIP(hipEventElapsedTime(&ms, d->ev0, d->ev1));
        d->last_forward_ms = ms;
        d->last_forward_voxels = V;
        d->last_output_voxels = V;
        return NSP_OK;
    } catch (const Error& e) {
        *d->err = e.msg;
        return e.code;
    }
}

int nsp_denoiser_output(const nsp_denoiser* d, float* output)
{
    if (!d || !output) return NSP_E_INVALID;
    try {
        NSP_HIP(hipSetDevice(d->device));
        NSP_HIP(hipMemcpyAsync(output, d->pred.p, sizeof(float) * d->last_output_voxels * 3, hipMemcpyDeviceToHost,
                               d->stream));
        NSP_HIP(hipStreamSynchronize(d->stream));
        return NSP_OK;
    } catch (const Error& e) {
        *d->err = e.msg;
        return e.code;
    }
}

int nsp_denoiser_last_forward(const nsp_denoiser* d, double* ms, double* flops)
{
    if (!d) return NSP_E_INVALID;
    if (ms) *ms = d->last_forward_ms;
    if (flops) {
        // multiply-adds on the real channels: 2 * 27 * (cin * cout) per voxel and layer
        double per = 0;
        for (int l = 0; l < DN_LAYERS; ++l) {
            const int ci = l == 0 ? d->cin_real : DN_WIDTH;
            const int co = l == DN_LAYERS - 1 ? (d->head == NSP_DENOISER_KERNEL ? 27 : 3) : DN_WIDTH;
            per += 2.0 * 27 * ci * co;
        }
        *flops = per * (double)d->last_forward_voxels;
    }
    return NSP_OK;
}


int nsp_denoiser_forward(nsp_denoiser* d, const float* input, int nb, int depth, int height, int width, float* output)
{
    if (!d || !input || !output || nb < 1 || depth < 1 || height < 1 || width < 1) return NSP_E_INVALID;
    try {
        NSP_HIP(hipSetDevice(d->device));
        Dims D = dims_of(nb, depth, height, width);
        const int64_t V = D.voxels();
        dn_alloc(*d, V);
        hipStream_t s = d->stream;
        DevBuf in;
        in.alloc(sizeof(float) * V * d->cin_real);
        NSP_HIP(hipMemcpyAsync(in.p, input, sizeof(float) * V * d->cin_real, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_pack_input, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, s, in.as<float>(), V,
                           d->cin_real, d->resid_ch, d->act[0].as<unsigned short>(), d->resid.as<float>());
        NSP_HIP(hipGetLastError());
        if (!d->ev0) NSP_HIP(hipEventCreate(&d->ev0));
        if (!d->ev1) NSP_HIP(hipEventCreate(&d->ev1));
        NSP_HIP(hipEventRecord(d->ev0, s));
        forward(*d, D);
        NSP_HIP(hipEventRecord(d->ev1, s));
        if (d->head == NSP_DENOISER_KERNEL)
            hipLaunchKernelGGL(k_kp_loss, dim3(1024), dim3(256), 0, s, d->out_f32.as<float>(), d->resid.as<float>(),
                               (const float*)nullptr, D, 1.0f, (unsigned short*)nullptr, d->pred.as<float>(),
                               (float*)nullptr);
        else
            hipLaunchKernelGGL(k_loss, dim3(1024), dim3(256), 0, s, d->out_f32.as<float>(), d->resid.as<float>(),
                               (const float*)nullptr, V, 1.0f, (unsigned short*)nullptr, d->pred.as<float>(),
                               (float*)nullptr);
        NSP_HIP(hipGetLastError());
        NSP_HIP(hipMemcpyAsync(output, d->pred.p, sizeof(float) * V * 3, hipMemcpyDeviceToHost, s));
        NSP_HIP(hipStreamSynchronize(s));
        float ms = 0;
        NSP_HIP(hipEventElapsedTime(&ms, d->ev0, d->ev1));
        d->last_forward_ms = ms;
        d->last_forward_voxels = V;
        d->last_output_voxels = V;
        return NSP_OK;
    } catch (const Error& e) {
        *d->err = e.msg;
        return e.code;
    }
}

int nsp_denoiser_loss_grad(nsp_denoiser* d, const float* input, const float* target, int nb, int depth, int height,
                           int width, float* loss, float* grad)
{
    if (!d || !input || !target || nb < 1 || depth < 1 || height < 1 || width < 1) return NSP_E_INVALID;
    try {
        NSP_HIP(hipSetDevice(d->device));
        Dims D = dims_of(nb, depth, height, width);
        const int64_t V = D.voxels();
        dn_alloc(*d, V);
        hipStream_t s = d->stream;
        DevBuf in;
        in.alloc(sizeof(float) * V * d->cin_real);
        NSP_HIP(hipMemcpyAsync(in.p, input, sizeof(float) * V * d->cin_real, hipMemcpyHostToDevice, s));
        NSP_HIP(hipMemcpyAsync(d->target.p, target, sizeof(float) * V * 3, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_pack_input, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, s, in.as<float>(), V,
                           d->cin_real, d->resid_ch, d->act[0].as<unsigned short>(), d->resid.as<float>());
        forward(*d, D);
        float l = loss_and_grad(*d, D, true);
        backward(*d, D, d->dy_a.as<unsigned short>());
        if (loss) *loss = l;
        // the context stream is non-blocking: copy on it, then wait
        if (grad) NSP_HIP(hipMemcpyAsync(grad, d->grads.p, sizeof(float) * d->nparams, hipMemcpyDeviceToHost, s));
        NSP_HIP(hipStreamSynchronize(s));
        return NSP_OK;
    } catch (const Error& e) {
        *d->err = e.msg;
        return e.code;
    }
}

int nsp_denoiser_train(nsp_denoiser* d, const float* inputs, const float* targets, int volumes, int depth, int height,
                       int width, int patch, int batch, int steps, float lr, uint64_t seed, float* losses)
{
    if (!d || !inputs || !targets || volumes < 1 || patch < 4 || batch < 1 || steps < 0) return NSP_E_INVALID;
    if (patch > depth || patch > height || patch > width) return NSP_E_INVALID;
    try {
        NSP_HIP(hipSetDevice(d->device));
        hipStream_t s = d->stream;
        const int64_t VV = (int64_t)volumes * depth * height * width;
        DevBuf vin, vtg;
        vin.alloc(sizeof(float) * VV * d->cin_real);
        vtg.alloc(sizeof(float) * VV * 3);
        NSP_HIP(hipMemcpyAsync(vin.p, inputs, sizeof(float) * VV * d->cin_real, hipMemcpyHostToDevice, s));
        NSP_HIP(hipMemcpyAsync(vtg.p, targets, sizeof(float) * VV * 3, hipMemcpyHostToDevice, s));
        Dims D = dims_of(batch, patch, patch, patch);
        const int64_t V = D.voxels();
        dn_alloc(*d, V);
        DevBuf dloss;
        dloss.alloc(sizeof(float) * std::max(steps, 1));
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        for (int st = 0; st < steps; ++st) {
            // the patches of a step are drawn by (seed, optimiser step count), so a
            // run resumed from a checkpoint draws what the unbroken run would
            hipLaunchKernelGGL(k_gather_patches, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, s, vin.as<float>(),
                               vtg.as<float>(), volumes, depth, height, width, d->cin_real, d->resid_ch, patch, batch,
                               seed, (int)d->adam_t, d->act[0].as<unsigned short>(), d->resid.as<float>(),
                               d->target.as<float>());
            NSP_HIP(hipGetLastError());
            forward(*d, D);
            loss_and_grad(*d, D, true, dloss.as<float>() + st);
            backward(*d, D, d->dy_a.as<unsigned short>());
            d->adam_t += 1;
            float c1 = 1.0f - std::pow(b1, (float)d->adam_t), c2 = 1.0f - std::pow(b2, (float)d->adam_t);
            hipLaunchKernelGGL(k_adam, dim3((unsigned)((d->nparams + 255) / 256)), dim3(256), 0, s,
                               d->params.as<float>(), d->grads.as<float>(), d->adam_m.as<float>(),
                               d->adam_v.as<float>(), d->nparams, lr, b1, b2, eps, c1, c2);
            NSP_HIP(hipGetLastError());
            pack_weights(*d);
        }
        if (losses && steps > 0)
            NSP_HIP(hipMemcpyAsync(losses, dloss.p, sizeof(float) * steps, hipMemcpyDeviceToHost, s));
        NSP_HIP(hipStreamSynchronize(s));
        return NSP_OK;
    } catch (const Error& e) {
        *d->err = e.msg;
        return e.code;
    }
}

// Stand-alone BF16 implicit-GEMM convolution for kernel-level tests:
// x [V][cin] f32 (cin in {16, 32}), w [cout][27][cin] f32, bias [cout] or NULL,
// y [V][cout] f32 (cout in {16, 32}).
int nsp_conv3d_bf16(nsp_context* ctx, const float* x, int nb, int depth, int height, int width, int cin,
                    const float* w, const float* bias, int cout, int relu, float* y)
{
    if (!ctx || !x || !w || !y) return NSP_E_INVALID;
    std::string* err = context_error(ctx);
    try {
        NSP_HIP(hipSetDevice(context_device(ctx)));
        if ((cin != 16 && cin != 32) || (cout != 16 && cout != 32)) raise(NSP_E_INVALID, "cin/cout must be 16 or 32");
        hipStream_t s = context_stream(ctx);
        Dims D = dims_of(nb, depth, height, width);
        const int64_t V = D.voxels();
        DevBuf dx, dxf, dw, dwf, dwb, db, dy;
        dxf.alloc(sizeof(float) * V * cin);
        dx.alloc(sizeof(unsigned short) * V * cin);
        dw.alloc(sizeof(float) * cout * 27 * cin);
        int kpf = kp_of(cin), kpb = kp_of(cout);
        dwf.alloc(sizeof(unsigned short) * cout * kpf);
        dwb.alloc(sizeof(unsigned short) * cin * kpb);
        dy.alloc(sizeof(float) * V * cout);
        NSP_HIP(hipMemcpyAsync(dxf.p, x, sizeof(float) * V * cin, hipMemcpyHostToDevice, s));
        NSP_HIP(hipMemcpyAsync(dw.p, w, sizeof(float) * cout * 27 * cin, hipMemcpyHostToDevice, s));
        if (bias) {
            db.alloc(sizeof(float) * cout);
            NSP_HIP(hipMemcpyAsync(db.p, bias, sizeof(float) * cout, hipMemcpyHostToDevice, s));
        }
        // f32 -> bf16 input (channels-last, cin already the padded count)
        {
            std::vector<unsigned short> hx((size_t)V * cin);
            for (int64_t i = 0; i < V * cin; ++i) {
                float f = x[i];
                uint32_t u;
                std::memcpy(&u, &f, 4);
                uint32_t r = ((u >> 16) & 1) + 0x7fff;  // round to nearest even
                hx[i] = (unsigned short)((u + r) >> 16);
            }
            NSP_HIP(hipMemcpyAsync(dx.p, hx.data(), sizeof(unsigned short) * hx.size(), hipMemcpyHostToDevice, s));
            NSP_HIP(hipStreamSynchronize(s));
        }
        int n = std::max(cout * kpf, cin * kpb);
        hipLaunchKernelGGL(k_pack_w, dim3((n + 255) / 256), dim3(256), 0, s, dw.as<float>(), cin, cout, kpf, kpb,
                           dwf.as<unsigned short>(), dwb.as<unsigned short>());
        conv(cin, cout, dx.as<unsigned short>(), D, dwf.as<unsigned short>(), bias ? db.as<float>() : nullptr, relu,
             nullptr, nullptr, dy.as<float>(), s);
        NSP_HIP(hipMemcpyAsync(y, dy.p, sizeof(float) * V * cout, hipMemcpyDeviceToHost, s));
        NSP_HIP(hipStreamSynchronize(s));
        return NSP_OK;
    } catch (const Error& e) {
        *err = e.msg;
        return e.code;
    }
}

}  // extern "C"
