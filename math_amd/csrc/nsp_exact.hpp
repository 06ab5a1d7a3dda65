// nsp_exact.hpp -- exact signed integers for the convex-hull predicates
// (host and device).  Facet normals are cofactor determinants of integer
// coordinate differences and a visibility test is their dot product with a
// difference; up to 4 dimensions these fit __int128, in 5 and 6 dimensions they
// need up to about 200 bits, held here in a 256-bit two's-complement integer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

namespace nsp {

struct I256 {
    uint64_t w[4];  // little-endian limbs, two's complement

    __host__ __device__ I256() : w{0, 0, 0, 0} {}
    __host__ __device__ I256(int64_t v)
    {
        const uint64_t s = v < 0 ? ~0ull : 0ull;
        w[0] = (uint64_t)v;
        w[1] = w[2] = w[3] = s;
    }

    __host__ __device__ friend I256 operator+(const I256& a, const I256& b)
    {
        I256 r;
        unsigned __int128 c = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            c += (unsigned __int128)a.w[i] + b.w[i];
            r.w[i] = (uint64_t)c;
            c >>= 64;
        }
        return r;
    }
    __host__ __device__ friend I256 operator-(const I256& a)
    {
        I256 r;
        unsigned __int128 c = 1;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            c += (unsigned __int128)(~a.w[i]);
            r.w[i] = (uint64_t)c;
            c >>= 64;
        }
        return r;
    }
    __host__ __device__ friend I256 operator-(const I256& a, const I256& b) { return a + (-b); }
    // low 256 bits of the product; exact whenever the product fits (the
    // callers' bounds keep every value below 2^200)
    __host__ __device__ friend I256 operator*(const I256& a, const I256& b)
    {
        uint64_t r[4] = {0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            unsigned __int128 c = 0;
#pragma unroll
            for (int j = 0; i + j < 4; ++j) {
                c += (unsigned __int128)a.w[i] * b.w[j] + r[i + j];
                r[i + j] = (uint64_t)c;
                c >>= 64;
            }
        }
        I256 o;
#pragma unroll
        for (int i = 0; i < 4; ++i) o.w[i] = r[i];
        return o;
    }
    __host__ __device__ I256& operator+=(const I256& b) { return *this = *this + b; }
    __host__ __device__ I256& operator-=(const I256& b) { return *this = *this - b; }

    __host__ __device__ int sign() const
    {
        if ((int64_t)w[3] < 0) return -1;
        return (w[0] | w[1] | w[2] | w[3]) ? 1 : 0;
    }
};

__host__ __device__ inline int sign_of(const I256& v) { return v.sign(); }
__host__ __device__ inline int sign_of(__int128 v) { return v < 0 ? -1 : (v > 0 ? 1 : 0); }

inline double to_double(__int128 v) { return (double)v; }
inline double to_double(const I256& v)
{
    const bool neg = v.sign() < 0;
    const I256 m = neg ? -v : v;
    double d = 0;
    for (int i = 3; i >= 0; --i) d = d * 18446744073709551616.0 + (double)m.w[i];
    return neg ? -d : d;
}

}  // namespace nsp
