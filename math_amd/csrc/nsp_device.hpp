// nsp_device.hpp -- device-side helpers shared by the n-space painter kernels
// (gfx950 / CDNA4; 64-lane wavefronts).
//
// Conventions follow the reference painter (templeblock/math):
//  * dot() is the fma chain of src/com/vec.h:258-267;
//  * the Philox4x32-10 stream layout is (pixel lo, sample, pass<<16|depth<<8|stream,
//    pixel hi) with the scene seed as key, so the FP64 CPU checker draws the
//    same uniforms for the same path;
//  * uniforms are 24-bit, (x >> 8) * 2^-24, exactly representable in float and double.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nsp {

constexpr int NMAX = 6;
constexpr int WAVE = 64;

template <int N, typename T>
struct Vec {
    T v[N];
    __host__ __device__ T& operator[](int i) { return v[i]; }
    __host__ __device__ const T& operator[](int i) const { return v[i]; }
};

template <int N, typename T>
__host__ __device__ __forceinline__ T dot(const T* a, const T* b)
{
    T r = a[0] * b[0];
#pragma unroll
    for (int i = 1; i < N; ++i) r = fma(a[i], b[i], r);
    return r;
}

template <int N, typename T>
__host__ __device__ __forceinline__ void normalize(T* v)
{
    T len = sqrt(dot<N, T>(v, v));
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = v[i] / len;
}

__device__ __forceinline__ void philox4x32(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                           uint32_t k1, uint32_t out[4])
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t lo0 = 0xD2511F53u * c0;
        uint32_t hi0 = __umulhi(0xD2511F53u, c0);
        uint32_t lo1 = 0xCD9E8D57u * c2;
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c2);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

template <typename T>
__device__ __forceinline__ T u01(uint32_t x)
{
    return (T)(x >> 8) * (T)(1.0 / 16777216.0);
}

template <typename T>
__device__ __forceinline__ void uniforms4(uint64_t seed, int64_t pixel, int sample, int pass, int depth, int stream,
                                          T u[4])
{
    uint32_t o[4];
    philox4x32((uint32_t)pixel, (uint32_t)sample, ((uint32_t)pass << 16) | ((uint32_t)depth << 8) | (uint32_t)stream,
               (uint32_t)((uint64_t)pixel >> 32), (uint32_t)seed, (uint32_t)(seed >> 32), o);
#pragma unroll
    for (int i = 0; i < 4; ++i) u[i] = u01<T>(o[i]);
}

__device__ __forceinline__ void sincospi_t(float x, float* s, float* c) { sincospif(x, s, c); }
__device__ __forceinline__ void sincospi_t(double x, double* s, double* c) { sincospi(x, s, c); }

// Determinant by cofactor expansion along the first row; for N=3,4 this is the
// explicit form of src/geometry/core/linear_algebra.h:194-236.  Rows 0..K-1 of
// m (row stride N), columns given by cols.
template <int N, int K, typename T>
struct Det {
    __host__ __device__ static T eval(const T* m, int row, const int* cols)
    {
        T s = 0;
#pragma unroll
        for (int j = 0; j < K; ++j) {
            int sub[K > 1 ? K - 1 : 1];
            int c = 0;
#pragma unroll
            for (int q = 0; q < K; ++q)
                if (q != j) sub[c++] = cols[q];
            T term = m[row * N + cols[j]] * Det<N, K - 1, T>::eval(m, row + 1, sub);
            if (j == 0)
                s = term;
            else if (j & 1)
                s = s - term;
            else
                s = s + term;
        }
        return s;
    }
};
template <int N, typename T>
struct Det<N, 2, T> {
    __host__ __device__ static T eval(const T* m, int row, const int* cols)
    {
        return m[row * N + cols[0]] * m[(row + 1) * N + cols[1]] - m[row * N + cols[1]] * m[(row + 1) * N + cols[0]];
    }
};
template <int N, typename T>
struct Det<N, 1, T> {
    __host__ __device__ static T eval(const T* m, int row, const int* cols) { return m[row * N + cols[0]]; }
};

// ortho_nn of N-1 vectors (rows of m), src/geometry/core/linear_algebra.h:167-186.
template <int N, typename T>
__host__ __device__ __forceinline__ void ortho_nn(const T* m, T* out)
{
#pragma unroll
    for (int i = 0; i < N; ++i) {
        int cols[N - 1];
        int c = 0;
#pragma unroll
        for (int q = 0; q < N; ++q)
            if (q != i) cols[c++] = q;
        T minor = Det<N, N - 1, T>::eval(m, 0, cols);
        out[i] = (i & 1) ? -minor : minor;
    }
}

// orthogonal_complement_of_unit_vector, src/geometry/core/complement.h:71-200
// (subspace construction for N <= 4, Gram-Schmidt for N >= 5).
template <int N, typename T>
__device__ __forceinline__ void orthogonal_complement(const T* u, T* basis /* (N-1)*N */)
{
    const T LIMIT = (T)0.1;
    if constexpr (N == 3) {
        T nc[3] = {0, 0, 0};
        if (fabs(u[0]) > LIMIT)
            nc[1] = 1;
        else
            nc[0] = 1;
        T e0[3], e1[3];
        e0[0] = +(u[1] * nc[2] - u[2] * nc[1]);
        e0[1] = -(u[0] * nc[2] - u[2] * nc[0]);
        e0[2] = +(u[0] * nc[1] - u[1] * nc[0]);
        normalize<3, T>(e0);
        e1[0] = +(u[1] * e0[2] - u[2] * e0[1]);
        e1[1] = -(u[0] * e0[2] - u[2] * e0[0]);
        e1[2] = +(u[0] * e0[1] - u[1] * e0[0]);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            basis[i] = e0[i];
            basis[3 + i] = e1[i];
        }
    } else if constexpr (N == 4) {
        int exclude = 0;
        for (; exclude < N - 2; ++exclude)
            if (fabs(u[exclude]) > LIMIT) break;
        T sb[(N - 1) * N];
#pragma unroll
        for (int q = 0; q < N; ++q) sb[(N - 2) * N + q] = u[q];
        int num = 0;
        for (int i = 0; num < N - 2; ++i)
            if (i != exclude) {
#pragma unroll
                for (int q = 0; q < N; ++q) sb[num * N + q] = (q == i) ? (T)1 : (T)0;
                ++num;
            }
        T t[N];
#pragma unroll
        for (int i = 0; i < N - 2; ++i) {
            ortho_nn<N, T>(sb, t);
            normalize<N, T>(t);
#pragma unroll
            for (int q = 0; q < N; ++q) sb[i * N + q] = t[q];
        }
        ortho_nn<N, T>(sb, t);
#pragma unroll
        for (int q = 0; q < N; ++q) sb[(N - 2) * N + q] = t[q];
#pragma unroll
        for (int k = 0; k < (N - 1) * N; ++k) basis[k] = sb[k];
    } else {
        int exclude = 0;
        for (; exclude < N - 1; ++exclude)
            if (fabs(u[exclude]) > LIMIT) break;
        T b[N][N], ob[N][N];
#pragma unroll
        for (int q = 0; q < N; ++q) b[0][q] = u[q];
        int num = 1;
        for (int i = 0; num < N; ++i)
            if (i != exclude) {
#pragma unroll
                for (int q = 0; q < N; ++q) b[num][q] = (q == i) ? (T)1 : (T)0;
                ++num;
            }
#pragma unroll
        for (int q = 0; q < N; ++q) ob[0][q] = b[0][q];
#pragma unroll
        for (int i = 1; i < N; ++i) {
            T sum[N];
#pragma unroll
            for (int q = 0; q < N; ++q) sum[q] = 0;
#pragma unroll
            for (int k = 0; k < i; ++k) {
                T d = dot<N, T>(b[i], ob[k]);
#pragma unroll
                for (int q = 0; q < N; ++q) sum[q] = sum[q] + d * ob[k][q];
            }
#pragma unroll
            for (int q = 0; q < N; ++q) ob[i][q] = b[i][q] - sum[q];
            normalize<N, T>(ob[i]);
        }
#pragma unroll
        for (int i = 0; i < N - 1; ++i)
#pragma unroll
            for (int q = 0; q < N; ++q) basis[i * N + q] = ob[i + 1][q];
    }
}

// random_cosine_weighted_on_hemisphere (src/painter/sampling/sphere.h:87-116)
// drawn in closed form: |tangent|^2 = u0, tangent direction uniform on S^(N-2).
template <int N, typename T>
__device__ __forceinline__ void cosine_hemisphere(const T* normal, const T* u, T* res)
{
    constexpr int M = N - 1;
    T dv[M];
    if constexpr (M == 2) {
        T s, c;
        sincospi_t((T)2 * u[1], &s, &c);
        dv[0] = c;
        dv[1] = s;
    } else if constexpr (M == 3) {
        T z = (T)1 - (T)2 * u[1];
        T rho = sqrt(fmax((T)0, (T)1 - z * z));
        T s, c;
        sincospi_t((T)2 * u[2], &s, &c);
        dv[0] = rho * c;
        dv[1] = rho * s;
        dv[2] = z;
    } else if constexpr (M == 4) {
        T a = u[1];
        T s1, c1, s2, c2;
        sincospi_t((T)2 * u[2], &s1, &c1);
        sincospi_t((T)2 * u[3], &s2, &c2);
        T ra = sqrt(a), rb = sqrt((T)1 - a);
        dv[0] = ra * c1;
        dv[1] = ra * s1;
        dv[2] = rb * c2;
        dv[3] = rb * s2;
    } else {
        T g[6];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            T r = sqrt((T)-2 * log((T)1 - u[1 + 2 * k]));
            T s, c;
            sincospi_t((T)2 * u[2 + 2 * k], &s, &c);
            g[2 * k] = r * c;
            g[2 * k + 1] = r * s;
        }
        T l2 = 0;
#pragma unroll
        for (int k = 0; k < M; ++k) l2 += g[k] * g[k];
        T l = sqrt(l2);
#pragma unroll
        for (int k = 0; k < M; ++k) dv[k] = g[k] / l;
    }
    T tl = sqrt(u[0]);
    T nc = sqrt((T)1 - u[0]);
    T basis[(N - 1) * N];
    orthogonal_complement<N, T>(normal, basis);
#pragma unroll
    for (int i = 0; i < N; ++i) res[i] = nc * normal[i];
#pragma unroll
    for (int k = 0; k < M; ++k) {
        T v = tl * dv[k];
#pragma unroll
        for (int i = 0; i < N; ++i) res[i] = res[i] + v * basis[k * N + i];
    }
}

// Wave-level compaction: returns this lane's slot in a queue whose tail counter
// is *tail, or -1 when pred is false.  One atomic per wave.
__device__ __forceinline__ int64_t wave_append(bool pred, unsigned long long* tail)
{
    uint64_t mask = __ballot(pred);
    if (mask == 0) return -1;
    int lane = __lane_id();
    int leader = __ffsll((unsigned long long)mask) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(tail, (unsigned long long)__popcll(mask));
    base = __shfl(base, leader, WAVE);
    if (!pred) return -1;
    uint64_t below = mask & ((lane == 0) ? 0ull : (~0ull >> (64 - lane)));
    return (int64_t)(base + __popcll(below));
}

// blockIdx remap so that consecutive logical blocks land on the same XCD
// (MI355X: 8 XCDs, workgroups dispatched round robin across them).
__device__ __forceinline__ unsigned xcd_block()
{
    unsigned nb = gridDim.x, b = blockIdx.x;
    if (nb % 8) return b;
    return (b % 8) * (nb / 8) + b / 8;
}

}  // namespace nsp
