// Device storage formats for gauge links and colour vectors on the even-odd
// checkerboard, with the accessors the kernels use.
//
// HBM layout:
//   links   : per (parity q, direction mu), blocks of 64 sites, each block an
//             array of 16/8/4-byte chunks x 64 lanes (see link_block_offset)
//   vectors : structure of arrays, v[off(r) + c*Vh + s], r = right-hand side,
//             c = colour, off(r) = r*rstride or a per-right-hand-side offset table
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
typedef float kq_v4f __attribute__((ext_vector_type(4)));
typedef unsigned int kq_v4u __attribute__((ext_vector_type(4)));
typedef unsigned int kq_v2u __attribute__((ext_vector_type(2)));
// 12-byte word (one dwordx3 load), 4-byte aligned: records of 12 bytes sit back to back
typedef unsigned int kq_v3u __attribute__((ext_vector_type(3), aligned(4)));
__device__ inline kq_v2d ld_nt(const kq_v2d *p) { return __builtin_nontemporal_load(p); }
__device__ inline kq_v4f ld_nt(const kq_v4f *p) { return __builtin_nontemporal_load(p); }
__device__ inline kq_v2f ld_nt(const kq_v2f *p) { return __builtin_nontemporal_load(p); }
__device__ inline kq_v4u ld_nt(const kq_v4u *p) { return __builtin_nontemporal_load(p); }
__device__ inline kq_v2u ld_nt(const kq_v2u *p) { return __builtin_nontemporal_load(p); }
__device__ inline kq_v3u ld_nt(const kq_v3u *p) { return __builtin_nontemporal_load(p); }
__device__ inline float ld_nt(const float *p) { return __builtin_nontemporal_load(p); }

// Streaming store of one complex number: the hopping term's output is not read
// again by the same launch, so its lines should not evict the neighbour colour
// vectors that other sites still gather from L2.
__device__ inline void st_nt(double2 *p, double2 v)
{
    const kq_v2d t = {v.x, v.y};
    __builtin_nontemporal_store(t, reinterpret_cast<kq_v2d *>(p));
}
__device__ inline void st_nt(float2 *p, float2 v)
{
    const kq_v2f t = {v.x, v.y};
    __builtin_nontemporal_store(t, reinterpret_cast<kq_v2f *>(p));
}
__device__ inline double2 ld_nt(const double2 *p)
{
    const kq_v2d v = __builtin_nontemporal_load(reinterpret_cast<const kq_v2d *>(p));
    return mk<double2>(v.x, v.y);
}
__device__ inline float2 ld_nt(const float2 *p)
{
    const kq_v2f v = __builtin_nontemporal_load(reinterpret_cast<const kq_v2f *>(p));
    return mk<float2>(v.x, v.y);
}

// two 16-bit fixed-point reals (low half: real part) -> floats in units of the
// stored scale: q / 32767 for |q| <= 32767
__device__ inline float2 q2f(unsigned int u)
{
    const float re = (float)(short)(u & 0xffffu), im = (float)(short)(u >> 16);
    return mk<float2>(re, im);
}

// ------------------------------------------------------------ links
// Link fields are stored per (parity q, direction mu) in blocks of 64 sites
// (one wavefront): block b of (q, mu) starts at byte ((q*4 + mu)*nb64 + b) *
// kBlock and holds, for each 16-, 8- or 4-byte chunk of a link, the 64 lanes'
// chunks back to back. A wavefront's load of one chunk is one contiguous,
// aligned 1 KiB / 512 B / 256 B segment, and the whole link of 64 sites is one
// contiguous region of 64 * kBytes.
__host__ __device__ inline size_t link_block_offset(int q, int mu, uint32_t s, uint32_t nb64, size_t block_bytes)
{
    return ((size_t)(q * 4 + mu) * nb64 + (s >> 6)) * block_bytes;
}

struct LinkD {  // FP64, 18 reals: 9 chunks of 16 B (one complex each)
    using R = double;
    static constexpr int kBytes = 144;
    static constexpr size_t kBlock = 64 * kBytes;
    const char *p;
    uint32_t nb64;
    // a link's stored words (fetch) and the matrix they encode (unpack); load = unpack(fetch)
    struct Raw {
        kq_v2d c[9];
    };
    __device__ inline Raw fetch(int q, int mu, uint32_t s) const
    {
        Raw r;
        const kq_v2d *b = reinterpret_cast<const kq_v2d *>(p + link_block_offset(q, mu, s, nb64, kBlock)) + (s & 63);
#pragma unroll
        for (int k = 0; k < 9; ++k)
            r.c[k] = ld_nt(b + 64 * k);
        return r;
    }
    __device__ inline void unpack(const Raw &r, double2 u[9]) const
    {
#pragma unroll
        for (int k = 0; k < 9; ++k)
            u[k] = mk<double2>(r.c[k].x, r.c[k].y);
    }
    __device__ inline void load(int q, int mu, uint32_t s, double2 u[9]) const { unpack(fetch(q, mu, s), u); }
};

// FP64 long links as 14 reals: rows 0 and 1 of an SU(3) matrix V and a complex scale w, U = w V
// with row 2 of V = conj(row0 x row1). A HISQ long link is c_Naik times a product of three U(3)
// links, so it has this form with w^3 = det U (canon_to_linkD14 checks every link and the
// library keeps LinkD otherwise). 7 chunks of 16 B: complex 0..5 of V, then w.
struct LinkD14 {
    using R = double;
    static constexpr int kBytes = 112;
    static constexpr size_t kBlock = 64 * kBytes;
    const char *p;
    uint32_t nb64;
    struct Raw {
        kq_v2d c[7];
    };
    __device__ inline Raw fetch(int q, int mu, uint32_t s) const
    {
        Raw r;
        const kq_v2d *b = reinterpret_cast<const kq_v2d *>(p + link_block_offset(q, mu, s, nb64, kBlock)) + (s & 63);
#pragma unroll
        for (int k = 0; k < 7; ++k)
            r.c[k] = ld_nt(b + 64 * k);
        return r;
    }
    __device__ inline void load(int q, int mu, uint32_t s, double2 u[9]) const { unpack(fetch(q, mu, s), u); }
    __device__ inline void unpack(const Raw &r, double2 u[9]) const
    {
        double2 v[6];
#pragma unroll
        for (int k = 0; k < 6; ++k)
            v[k] = mk<double2>(r.c[k].x, r.c[k].y);
        const kq_v2d w = r.c[6];
        // row 2 of V = conj(row0 x row1), then U = w V
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
            const double2 a1 = v[c1], b2 = v[3 + c2], a2 = v[c2], b1 = v[3 + c1];
            const double re = (a1.x * b2.x - a1.y * b2.y) - (a2.x * b1.x - a2.y * b1.y);
            const double im = (a1.x * b2.y + a1.y * b2.x) - (a2.x * b1.y + a2.y * b1.x);
            u[6 + c] = mk<double2>(w.x * re + w.y * im, w.y * re - w.x * im);  // w * conj(re + i im)
        }
#pragma unroll
        for (int k = 0; k < 6; ++k)
            u[k] = mk<double2>(w.x * v[k].x - w.y * v[k].y, w.x * v[k].y + w.y * v[k].x);
    }
};

struct LinkF {  // FP32, 18 reals: 4 chunks of 16 B (complex 0..7) + 1 of 8 B (complex 8)
    using R = float;
    static constexpr int kBytes = 72;
    static constexpr size_t kBlock = 64 * kBytes;
    const char *p;
    uint32_t nb64;
    struct Raw {
        kq_v4f c[4];
        kq_v2f d;
    };
    __device__ inline Raw fetch(int q, int mu, uint32_t s) const
    {
        Raw r;
        const char *b = p + link_block_offset(q, mu, s, nb64, kBlock);
        const uint32_t l = s & 63;
#pragma unroll
        for (int c = 0; c < 4; ++c)
            r.c[c] = ld_nt(reinterpret_cast<const kq_v4f *>(b + 1024 * c) + l);
        r.d = ld_nt(reinterpret_cast<const kq_v2f *>(b + 4096) + l);
        return r;
    }
    __device__ inline void unpack(const Raw &r, float2 u[9]) const
    {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            u[2 * c] = mk<float2>(r.c[c].x, r.c[c].y);
            u[2 * c + 1] = mk<float2>(r.c[c].z, r.c[c].w);
        }
        u[8] = mk<float2>(r.d.x, r.d.y);
    }
    __device__ inline void load(int q, int mu, uint32_t s, float2 u[9]) const { unpack(fetch(q, mu, s), u); }
};

// 16-bit links, 18 reals: each real is a 16-bit fixed-point integer q with
// |q| <= 32767 and U = a * q, where a = max|entry| / 32767 is one FP32 scale per
// link (the lattice "half" format: 15 bits per entry relative to the largest).
// Chunks: 16 B (complex 0..3), 16 B (complex 4..7), 8 B (complex 8 + scale).
struct LinkH {
    using R = float;
    static constexpr int kBytes = 40;
    static constexpr size_t kBlock = 64 * kBytes;
    const char *p;
    uint32_t nb64;
    struct Raw {
        kq_v4u c0, c1;
        kq_v2u c2;
    };
    __device__ inline Raw fetch(int q, int mu, uint32_t s) const
    {
        Raw r;
        const char *b = p + link_block_offset(q, mu, s, nb64, kBlock);
        const uint32_t l = s & 63;
        r.c0 = ld_nt(reinterpret_cast<const kq_v4u *>(b) + l);
        r.c1 = ld_nt(reinterpret_cast<const kq_v4u *>(b + 1024) + l);
        r.c2 = ld_nt(reinterpret_cast<const kq_v2u *>(b + 2048) + l);
        return r;
    }
    __device__ inline void load(int q, int mu, uint32_t s, float2 u[9]) const { unpack(fetch(q, mu, s), u); }
    __device__ inline void unpack(const Raw &r, float2 u[9]) const
    {
        const kq_v4u c0 = r.c0, c1 = r.c1;
        const kq_v2u c2 = r.c2;
        const float a = __uint_as_float(c2.y);
        const unsigned int h[9] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w, c2.x};
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const float2 f = q2f(h[k]);
            u[k] = mk<float2>(f.x * a, f.y * a);
        }
    }
};

// 16-bit links, 12 reals (rows 0 and 1 of an SU(3) matrix V as fixed point
// q / 32767), one signed FP32 scale b per link: U = b * V with row 2 of V =
// conj(row0 x row1). Chunks: 16 B (complex 0..3), then one 12 B record of complex 4, 5 and the
// scale (its bits), so a link takes two loads.
struct LinkH12 {
    using R = float;
    static constexpr int kBytes = 28;
    static constexpr size_t kBlock = 64 * kBytes;
    const char *p;
    uint32_t nb64;
    struct Raw {
        kq_v4u c0;
        kq_v3u c1;
    };
    __device__ inline Raw fetch(int q, int mu, uint32_t s) const
    {
        Raw r;
        const char *b = p + link_block_offset(q, mu, s, nb64, kBlock);
        const uint32_t l = s & 63;
        r.c0 = ld_nt(reinterpret_cast<const kq_v4u *>(b) + l);
        r.c1 = ld_nt(reinterpret_cast<const kq_v3u *>(b + 1024 + 12 * l));
        return r;
    }
    __device__ inline void load(int q, int mu, uint32_t s, float2 u[9]) const { unpack(fetch(q, mu, s), u); }
    __device__ inline void unpack(const Raw &r, float2 u[9]) const
    {
        const kq_v4u c0 = r.c0;
        const kq_v3u c1 = r.c1;
        const float a = __uint_as_float(c1.z);
        constexpr float kQ = 1.0f / 32767.0f;
        float2 v[6];
        v[0] = q2f(c0.x);
        v[1] = q2f(c0.y);
        v[2] = q2f(c0.z);
        v[3] = q2f(c0.w);
        v[4] = q2f(c1.x);
        v[5] = q2f(c1.y);
#pragma unroll
        for (int k = 0; k < 6; ++k)
            v[k] = mk<float2>(v[k].x * kQ, v[k].y * kQ);
        // row2 = conj(row0 x row1)
        float2 w[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int c1i = (c + 1) % 3, c2i = (c + 2) % 3;
            const float2 a1 = v[c1i], b2 = v[3 + c2i], a2 = v[c2i], b1 = v[3 + c1i];
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

// FP32, 12 reals (rows 0 and 1 of an SU(3) matrix V), one signed FP32 scale b
// per link: U = b * V with row 2 of V = conj(row0 x row1).
// Chunks: 3 x 16 B (complex 0..5), 4 B (scale).
struct LinkF12 {
    using R = float;
    static constexpr int kBytes = 52;
    static constexpr size_t kBlock = 64 * kBytes;
    const char *p;
    uint32_t nb64;
    struct Raw {
        kq_v4f c[3];
        float a;
    };
    __device__ inline Raw fetch(int q, int mu, uint32_t s) const
    {
        Raw r;
        const char *b = p + link_block_offset(q, mu, s, nb64, kBlock);
        const uint32_t l = s & 63;
#pragma unroll
        for (int c = 0; c < 3; ++c)
            r.c[c] = ld_nt(reinterpret_cast<const kq_v4f *>(b + 1024 * c) + l);
        r.a = ld_nt(reinterpret_cast<const float *>(b + 3072) + l);
        return r;
    }
    __device__ inline void load(int q, int mu, uint32_t s, float2 u[9]) const { unpack(fetch(q, mu, s), u); }
    __device__ inline void unpack(const Raw &r, float2 u[9]) const
    {
        float2 v[6];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            v[2 * c] = mk<float2>(r.c[c].x, r.c[c].y);
            v[2 * c + 1] = mk<float2>(r.c[c].z, r.c[c].w);
        }
        const float a = r.a;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int c1i = (c + 1) % 3, c2i = (c + 2) % 3;
            const float2 a1 = v[c1i], b2 = v[3 + c2i], a2 = v[c2i], b1 = v[3 + c1i];
            const float re = (a1.x * b2.x - a1.y * b2.y) - (a2.x * b1.x - a2.y * b1.y);
            const float im = (a1.x * b2.y + a1.y * b2.x) - (a2.x * b1.y + a2.y * b1.x);
            u[6 + c] = mk<float2>(re * a, -im * a);
        }
#pragma unroll
        for (int k = 0; k < 6; ++k)
            u[k] = mk<float2>(v[k].x * a, v[k].y * a);
    }
};

// Row 2 of an SU(3) matrix from rows 0 and 1 (v[0..2], v[3..5]): conj(row0 x row1), times the
// complex scale w: u[6 + c] = w conj(row0 x row1)_c
template <typename R>
__device__ inline void su3_row2_scaled(const c2_t<R> v[6], R wx, R wy, c2_t<R> u[9])
{
    using V2 = c2_t<R>;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
        const V2 a1 = v[c1], b2 = v[3 + c2], a2 = v[c2], b1 = v[3 + c1];
        const R re = (a1.x * b2.x - a1.y * b2.y) - (a2.x * b1.x - a2.y * b1.y);
        const R im = (a1.x * b2.y + a1.y * b2.x) - (a2.x * b1.y + a2.y * b1.x);
        u[6 + c] = mk<V2>(wx * re + wy * im, wy * re - wx * im);
    }
}

// FP32 long links that are complex (not real) multiples of SU(3) matrices: a HISQ long link is
// c_Naik times a product of three U(3) links, whose determinant carries a phase. Rows 0 and 1 of
// V in SU(3) and w with w^3 = det U, U = w V. Chunks: 3 x 16 B (complex 0..5), 8 B (w): 56 B.
struct LinkF12C {
    using R = float;
    static constexpr int kBytes = 56;
    static constexpr size_t kBlock = 64 * kBytes;
    const char *p;
    uint32_t nb64;
    struct Raw {
        kq_v4f c[3];
        kq_v2f w;
    };
    __device__ inline Raw fetch(int q, int mu, uint32_t s) const
    {
        Raw r;
        const char *b = p + link_block_offset(q, mu, s, nb64, kBlock);
        const uint32_t l = s & 63;
#pragma unroll
        for (int c = 0; c < 3; ++c)
            r.c[c] = ld_nt(reinterpret_cast<const kq_v4f *>(b + 1024 * c) + l);
        r.w = ld_nt(reinterpret_cast<const kq_v2f *>(b + 3072) + l);
        return r;
    }
    __device__ inline void load(int q, int mu, uint32_t s, float2 u[9]) const { unpack(fetch(q, mu, s), u); }
    __device__ inline void unpack(const Raw &r, float2 u[9]) const
    {
        float2 v[6];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            v[2 * c] = mk<float2>(r.c[c].x, r.c[c].y);
            v[2 * c + 1] = mk<float2>(r.c[c].z, r.c[c].w);
        }
        const kq_v2f w = r.w;
        su3_row2_scaled<float>(v, w.x, w.y, u);
#pragma unroll
        for (int k = 0; k < 6; ++k)
            u[k] = mk<float2>(w.x * v[k].x - w.y * v[k].y, w.x * v[k].y + w.y * v[k].x);
    }
};

// The same in 16 bits: rows 0 and 1 of V as fixed point q / 32767 and w in FP32.
// Chunks: 16 B (complex 0..3), 16 B (complex 4, 5 and the bits of w): 32 B in two loads.
struct LinkH12C {
    using R = float;
    static constexpr int kBytes = 32;
    static constexpr size_t kBlock = 64 * kBytes;
    const char *p;
    uint32_t nb64;
    struct Raw {
        kq_v4u c0;
        kq_v4u c1;
    };
    __device__ inline Raw fetch(int q, int mu, uint32_t s) const
    {
        Raw r;
        const char *b = p + link_block_offset(q, mu, s, nb64, kBlock);
        const uint32_t l = s & 63;
        r.c0 = ld_nt(reinterpret_cast<const kq_v4u *>(b) + l);
        r.c1 = ld_nt(reinterpret_cast<const kq_v4u *>(b + 1024) + l);
        return r;
    }
    __device__ inline void load(int q, int mu, uint32_t s, float2 u[9]) const { unpack(fetch(q, mu, s), u); }
    __device__ inline void unpack(const Raw &r, float2 u[9]) const
    {
        const kq_v4u c0 = r.c0;
        const kq_v4u c1 = r.c1;
        const kq_v2f w = {__uint_as_float(c1.z), __uint_as_float(c1.w)};
        constexpr float kQ = 1.0f / 32767.0f;
        const unsigned int h[6] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y};
        float2 v[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const float2 f = q2f(h[k]);
            v[k] = mk<float2>(f.x * kQ, f.y * kQ);
        }
        su3_row2_scaled<float>(v, w.x, w.y, u);
#pragma unroll
        for (int k = 0; k < 6; ++k)
            u[k] = mk<float2>(w.x * v[k].x - w.y * v[k].y, w.x * v[k].y + w.y * v[k].x);
    }
};

inline uint32_t nb64_of(uint32_t Vh) { return (Vh + 63) / 64; }
// bytes of one link field (both parities, 4 directions) in format LK
template <class LK> inline size_t link_field_bytes(uint32_t Vh) { return (size_t)8 * nb64_of(Vh) * LK::kBlock; }

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
// Each complex product is accumulated by four chained fused multiply-adds into the running sum
// (acc = fma(a, b, acc)), not formed apart and then added: 36 instead of 54 FP operations per
// matrix-vector product. The batched hopping terms issue NR such products per hop and are
// bound by them as much as by memory (four right-hand sides in FP32: ~4400 FP instructions
// per wave against ~2900 with the chains).
__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }

// acc += U v
template <typename V2, typename R>
__device__ inline void umv_add(const V2 u[9], const V2 v[3], R acc[6])
{
#pragma unroll
    for (int row = 0; row < 3; ++row) {
#pragma unroll
        for (int col = 0; col < 3; ++col) {
            const V2 a = u[3 * row + col], b = v[col];
            acc[2 * row] = fmadd((R)a.x, (R)b.x, acc[2 * row]);
            acc[2 * row] = fmadd(-(R)a.y, (R)b.y, acc[2 * row]);
            acc[2 * row + 1] = fmadd((R)a.x, (R)b.y, acc[2 * row + 1]);
            acc[2 * row + 1] = fmadd((R)a.y, (R)b.x, acc[2 * row + 1]);
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
            acc[2 * row] = fmadd(-(R)a.x, (R)b.x, acc[2 * row]);
            acc[2 * row] = fmadd(-(R)a.y, (R)b.y, acc[2 * row]);
            acc[2 * row + 1] = fmadd(-(R)a.x, (R)b.y, acc[2 * row + 1]);
            acc[2 * row + 1] = fmadd((R)a.y, (R)b.x, acc[2 * row + 1]);
        }
    }
}

}  // namespace kq
