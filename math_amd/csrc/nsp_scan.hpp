// nsp_scan.hpp -- device-wide exclusive scans for the tree build (u64 values).
//
// A scan runs in three phases: per-tile reduction, a (recursive) scan of the
// tile sums, and a per-tile down-sweep that hands each element its exclusive
// prefix.  The element values and the consumer are functors, so the tree
// build never materialises its (child, pair) flag array: the flags are
// recomputed in the down-sweep and consumed in place.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nsp_internal.hpp"

namespace nsp {

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int64_t SCAN_TILE = (int64_t)SCAN_BLOCK * SCAN_ITEMS;

__device__ __forceinline__ uint64_t wave_inclusive_scan(uint64_t v)
{
    int lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint64_t o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

// Block-wide exclusive scan of one value per thread (blockDim == SCAN_BLOCK).
// Returns the exclusive prefix; *total receives the block sum.
__device__ __forceinline__ uint64_t block_exclusive_scan(uint64_t v, uint64_t* total)
{
    __shared__ uint64_t wsum[SCAN_BLOCK / 64];
    int lane = __lane_id();
    int w = threadIdx.x / 64;
    uint64_t inc = wave_inclusive_scan(v);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint64_t off = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < SCAN_BLOCK / 64; ++k) {
        if (k < w) off += wsum[k];
        tot += wsum[k];
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

__device__ __forceinline__ uint64_t block_reduce(uint64_t v)
{
    uint64_t t;
    (void)block_exclusive_scan(v, &t);
    return t;
}

template <class Op>
__global__ void __launch_bounds__(SCAN_BLOCK) k_tile_reduce(Op op, int64_t n, uint64_t* tile_sums)
{
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    uint64_t s = 0;
#pragma unroll 4
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        int64_t i = base + (int64_t)k * SCAN_BLOCK + threadIdx.x;
        if (i < n) s += op.value(i);
    }
    uint64_t t = block_reduce(s);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = t;
}

template <class Op>
__global__ void __launch_bounds__(SCAN_BLOCK) k_tile_downsweep(Op op, int64_t n, const uint64_t* tile_offsets)
{
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    uint64_t carry = tile_offsets[blockIdx.x];
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        int64_t i = base + (int64_t)k * SCAN_BLOCK + threadIdx.x;
        uint64_t v = (i < n) ? op.value(i) : 0;
        uint64_t tot;
        uint64_t ex = block_exclusive_scan(v, &tot);
        if (i < n) op.emit(i, carry + ex, v);
        carry += tot;
    }
}

// Materialised exclusive scan of a u64 array (used for tile sums and small arrays).
struct ArrayScanOp {
    const uint64_t* in;
    uint64_t* out;
    __device__ uint64_t value(int64_t i) const { return in[i]; }
    __device__ void emit(int64_t i, uint64_t ex, uint64_t) const { out[i] = ex; }
};

void scan_u64(const uint64_t* d_in, uint64_t* d_out, int64_t n, uint64_t* total_host, hipStream_t s);

// Runs the three phases for a functor scan over n elements; returns the total.
template <class Op>
uint64_t scan_functor(const Op& op, int64_t n, hipStream_t s)
{
    if (n <= 0) return 0;
    int64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    DevBuf sums, offs;
    sums.alloc(sizeof(uint64_t) * tiles);
    offs.alloc(sizeof(uint64_t) * tiles);
    hipLaunchKernelGGL(k_tile_reduce<Op>, dim3((unsigned)tiles), dim3(SCAN_BLOCK), 0, s, op, n, sums.as<uint64_t>());
    NSP_HIP(hipGetLastError());
    uint64_t total = 0;
    scan_u64(sums.as<uint64_t>(), offs.as<uint64_t>(), tiles, &total, s);
    hipLaunchKernelGGL(k_tile_downsweep<Op>, dim3((unsigned)tiles), dim3(SCAN_BLOCK), 0, s, op, n,
                       offs.as<uint64_t>());
    NSP_HIP(hipGetLastError());
    NSP_HIP(hipStreamSynchronize(s));
    return total;
}

}  // namespace nsp
