// Device storage formats for gauge links and colour vectors on the even-odd
// checkerboard, with the accessors the kernels use.
//
// HBM layout (structure of arrays, site index fastest so that a wavefront's
// 64 lanes read 64 consecutive sites with one coalesced load per component):
//   links   : comp[((q*4 + mu)*NC + k)*Vh + s]   q = parity block, NC components
//   scales  : scale[(q*4 + mu)*Vh + s]           (FP16 formats only)
//   vectors : v[off(r) + c*Vh + s]               r = right-hand side, c = colour,
//             off(r) = r*rstride or a per-right-hand-side offset table
//
// Device links are pre-scaled at load time: fat links by 1/2 and long links by
// c3000 = -1/48 (dirac_hisq.cpp:40, 473, 480), so the hopping term is
//   D psi(x) = sum_mu U'(x) psi(x+mu) + N'(x) psi(x+3mu) - U'(x-mu)^+ psi(x-mu) - N'(x-3mu)^+ psi(x-3mu).
#pragma once

#include "kq_device.h"

namespace kq {

// Streaming (non-temporal) loads: each link is read once per hopping-term
// application, so it should not displace the reused colour vectors in L2.
typedef double kq_v2d __attribute__((ext_vector_type(2)));
typedef float kq_v2f __attribute__((ext_vector_type(2)));
__device__ inline double2 ld_stream(const double2 *p)
{
    kq_v2d v = __builtin_nontemporal_load(reinterpret_cast<const kq_v2d *>(p));
    return mk<double2>(v.x, v.y);
}
__device__ inline float2 ld_stream(const float2 *p)
{
    kq_v2f v = __builtin_nontemporal_load(reinterpret_cast<const kq_v2f *>(p));
    return mk<float2>(v.x, v.y);
}
__device__ inline __half2 ld_stream(const __half2 *p)
{
    unsigned int u = __builtin_nontemporal_load(reinterpret_cast<const unsigned int *>(p));
    return *reinterpret_cast<__half2 *>(&u);
}
__device__ inline float ld_stream(const float *p)
{
    return __builtin_nontemporal_load(p);
}

// ------------------------------------------------------------ links
struct LinkD {  // FP64, 18 reals
    using R = double;
    static constexpr int kBytes = 144;
    const double2 *p;
    uint32_t Vh;
    __device__ inline void load(int q, int mu, uint32_t s, double2 u[9]) const
    {
        const double2 *b = p + (size_t)((q * 4 + mu) * 9) * Vh + s;
#pragma unroll
        for (int k = 0; k < 9; ++k)
            u[k] = ld_stream(b + (size_t)k * Vh);
    }
};

struct LinkF {  // FP32, 18 reals
    using R = float;
    static constexpr int kBytes = 72;
    const float2 *p;
    uint32_t Vh;
    __device__ inline void load(int q, int mu, uint32_t s, float2 u[9]) const
    {
        const float2 *b = p + (size_t)((q * 4 + mu) * 9) * Vh + s;
#pragma unroll
        for (int k = 0; k < 9; ++k)
            u[k] = ld_stream(b + (size_t)k * Vh);
    }
};

// FP16, 18 reals, one FP32 scale per link: U = scale * half(U/scale).
struct LinkH {
    using R = float;
    static constexpr int kBytes = 40;
    const __half2 *p;
    const float *scale;
    uint32_t Vh;
    __device__ inline void load(int q, int mu, uint32_t s, float2 u[9]) const
    {
        const __half2 *b = p + (size_t)((q * 4 + mu) * 9) * Vh + s;
        const float a = ld_stream(scale + (size_t)(q * 4 + mu) * Vh + s);
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            __half2 h = ld_stream(b + (size_t)k * Vh);
            float2 f = __half22float2(h);
            u[k] = mk<float2>(f.x * a, f.y * a);
        }
    }
};

// FP16, 12 reals (rows 0 and 1 of an SU(3) matrix V), one signed FP32 scale b
// per link: U = b * V with row 2 of V = conj(row0 x row1).
struct LinkH12 {
    using R = float;
    static constexpr int kBytes = 28;
    const __half2 *p;
    const float *scale;
    uint32_t Vh;
    __device__ inline void load(int q, int mu, uint32_t s, float2 u[9]) const
    {
        const __half2 *b = p + (size_t)((q * 4 + mu) * 6) * Vh + s;
        const float a = ld_stream(scale + (size_t)(q * 4 + mu) * Vh + s);
        float2 v[6];
#pragma unroll
        for (int k = 0; k < 6; ++k)
            v[k] = __half22float2(ld_stream(b + (size_t)k * Vh));
        // row2 = conj(row0 x row1)
        float2 w[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
            const float2 a1 = v[c1], b2 = v[3 + c2], a2 = v[c2], b1 = v[3 + c1];
            const float re = (a1.x * b2.x - a1.y * b2.y) - (a2.x * b1.x - a2.y * b1.y);
            const float im = (a1.x * b2.y + a1.y * b2.x) - (a2.x * b1.y + a2.y * b1.x);
            w[c] = mk<float2>(re, -im);
        }
#pragma unroll
        for (int k = 0; k < 6; ++k)
            u[k] = mk<float2>(v[k].x * a, v[k].y * a);
#pragma unroll
        for (int c = 0; c < 3; ++c)
            u[6 + c] = mk<float2>(w[c].x * a, w[c].y * a);
    }
};

// ------------------------------------------------------------ colour vectors
template <typename R> struct Vec {
    using V2 = c2_t<R>;
    V2 *p;
    uint32_t Vh;
    size_t rstride;              // elements between consecutive right-hand sides (3*Vh when packed)
    const size_t *roff = nullptr;  // per right-hand-side element offsets (device); overrides rstride
    __device__ inline size_t base(int r) const { return roff ? roff[r] : (size_t)r * rstride; }
    __device__ inline void load(int r, uint32_t s, V2 v[3]) const
    {
        const V2 *b = p + base(r) + s;
#pragma unroll
        for (int c = 0; c < 3; ++c)
            v[c] = b[(size_t)c * Vh];
    }
    __device__ inline void store(int r, uint32_t s, const V2 v[3]) const
    {
        V2 *b = p + base(r) + s;
#pragma unroll
        for (int c = 0; c < 3; ++c)
            b[(size_t)c * Vh] = v[c];
    }
};

// ------------------------------------------------------------ SU(3) x vector
// acc += U v
template <typename V2, typename R>
__device__ inline void umv_add(const V2 u[9], const V2 v[3], R acc[6])
{
#pragma unroll
    for (int row = 0; row < 3; ++row) {
#pragma unroll
        for (int col = 0; col < 3; ++col) {
            const V2 a = u[3 * row + col], b = v[col];
            acc[2 * row] += (R)a.x * (R)b.x - (R)a.y * (R)b.y;
            acc[2 * row + 1] += (R)a.x * (R)b.y + (R)a.y * (R)b.x;
        }
    }
}

// acc -= U^+ v
template <typename V2, typename R>
__device__ inline void udmv_sub(const V2 u[9], const V2 v[3], R acc[6])
{
#pragma unroll
    for (int row = 0; row < 3; ++row) {
#pragma unroll
        for (int col = 0; col < 3; ++col) {
            const V2 a = u[3 * col + row], b = v[col];
            acc[2 * row] -= (R)a.x * (R)b.x + (R)a.y * (R)b.y;
            acc[2 * row + 1] -= (R)a.x * (R)b.y - (R)a.y * (R)b.x;
        }
    }
}

}  // namespace kq
