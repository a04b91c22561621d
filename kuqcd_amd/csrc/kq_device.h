// Device-side building blocks shared by every kernel of the HISQ multishift
// solver: HIP error checks, complex arithmetic on hip vector types, fast
// integer division, checkerboard geometry with in-kernel periodic wrap, and
// wavefront/workgroup reductions.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

#define KQ_HIP_CHECK(expr)                                                                           \
    do {                                                                                             \
        hipError_t kq_err_ = (expr);                                                                 \
        if (kq_err_ != hipSuccess) {                                                                 \
            throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(kq_err_) + " at " + \
                                     __FILE__ + ":" + std::to_string(__LINE__) + ": " #expr);        \
        }                                                                                            \
    } while (0)

namespace kq {

constexpr int kWave = 64;

// ---------------------------------------------------------------- complex
template <typename R> struct C2;
template <> struct C2<double> { using T = double2; };
template <> struct C2<float> { using T = float2; };
template <typename R> using c2_t = typename C2<R>::T;

template <typename V> __host__ __device__ inline V mk(decltype(V::x) a, decltype(V::x) b)
{
    V v;
    v.x = a;
    v.y = b;
    return v;
}

// ---------------------------------------------------------------- fast division
// Unsigned 32-bit division by a runtime constant (round-up multiplier method,
// Granlund & Montgomery 1994): q = (umulhi(n, m) + n) >> l with
// l = ceil(log2 d) and m = floor(2^32 (2^l - d) / d) + 1. Exact for n < 2^31.
struct FastDiv {
    uint32_t d = 1, m = 1, l = 0;
    FastDiv() = default;
    __host__ FastDiv(uint32_t d_) : d(d_)
    {
        l = 0;
        while ((1ull << l) < d)
            ++l;
        m = (uint32_t)((((1ull << 32) * ((1ull << l) - d)) / d) + 1);
    }
    __host__ __device__ inline uint32_t div(uint32_t n) const
    {
#ifdef __HIP_DEVICE_COMPILE__
        uint32_t t = __umulhi(n, m);
#else
        uint32_t t = (uint32_t)(((uint64_t)n * m) >> 32);
#endif
        return (t + n) >> l;
    }
};

// ---------------------------------------------------------------- geometry
// Checkerboard geometry of an Lx x Ly x Lz x Lt periodic lattice.
// Site s of parity q: s = xcb + (Lx/2)*(y + Ly*(z + Lz*t)), full x = 2*xcb + ((y+z+t+q)&1).
// Neighbour at hop h along mu: the checkerboard index of the wrapped full
// coordinate, (x+h mod Lx)>>1 along x (index_utils.cpp:91-111,
// space_time.cpp:187-211, dirac_hisq.cpp:139-231). Valid for any extent >= 1.
struct Geom {
    int L[4];
    int lxh;
    uint32_t Vh;
    uint32_t stride[4];  // checkerboard index stride of y, z, t (stride[0] unused)
    FastDiv dx, dy, dz;  // divisions by Lx/2, Ly, Lz
    // lattice tiles of the hopping term's block order (see tile_site): the 8 XCDs split y into
    // 2^lyp ranges of yc rows, z into 2^lzp ranges of zc planes and t into the rest, ranges of tc;
    // slab = (Lx/2)*yc sites of one (x, y) range, plane = (Lx/2)*Ly
    int tile;
    uint32_t plane, slab, yc, zc, tc, lyp, lzp;
    FastDiv dslab, dzc;
};

// Largest (x, y) slab of a tile, in sites of one parity: the 32^4 plane (512 sites) keeps the
// gathered vectors of the planes an XCD has in flight within its L2; the 1152-site planes of
// 48^3x96 did not (FP64 term at 74 % of HBM peak there against 83 % at 32^4), so such planes
// are split along y between XCDs.
constexpr uint32_t kTileSlab = 640;
// KQ_TILE_SLAB=<sites> replaces kTileSlab (a tuning switch for the hopping terms' site order;
// results are unchanged except for the grouping of FP64 partial sums)
inline uint32_t tile_slab_limit()
{
    static const uint32_t v = [] {
        const char *e = getenv("KQ_TILE_SLAB");
        const long n = e ? atol(e) : 0;
        return n > 0 ? (uint32_t)n : kTileSlab;
    }();
    return v;
}

// Tile split of the hopping terms' site order (tile_site): 8 tiles of (y ranges, z ranges,
// t ranges), the first split that divides the lattice with an (x, y) slab of at most slab_limit
// sites of one parity, else the first that divides it
__host__ inline void set_tiles(Geom &g, uint64_t slab_limit)
{
    const int *L = g.L;
    static const int kSplit[3][3] = {{1, 4, 2}, {2, 4, 1}, {4, 2, 1}};
    int pick = -1;
    for (int i = 0; i < 3; ++i) {
        const int yp = kSplit[i][0], zp = kSplit[i][1], tp = kSplit[i][2];
        if (L[1] % yp != 0 || L[2] % zp != 0 || L[3] % tp != 0)
            continue;
        const bool fits = (uint64_t)g.lxh * (L[1] / yp) <= slab_limit;
        if (pick < 0 || fits) {
            pick = i;
            if (fits)
                break;
        }
    }
    g.tile = pick >= 0 ? 1 : 0;
    const int yp = pick >= 0 ? kSplit[pick][0] : 1, zp = pick >= 0 ? kSplit[pick][1] : 4;
    const int tp = pick >= 0 ? kSplit[pick][2] : 2;
    g.lyp = yp == 4 ? 2 : yp == 2 ? 1 : 0;
    g.lzp = zp == 4 ? 2 : 1;
    g.plane = g.stride[2];
    g.yc = (uint32_t)(L[1] / yp);
    g.zc = (uint32_t)(L[2] / zp);
    g.tc = (uint32_t)(L[3] / tp);
    g.slab = (uint32_t)g.lxh * g.yc;
    g.dslab = FastDiv(g.slab > 0 ? g.slab : 1);
    g.dzc = FastDiv(g.zc > 0 ? g.zc : 1);
}

__host__ inline Geom make_geom(const int L[4])
{
    Geom g;
    for (int i = 0; i < 4; ++i)
        g.L[i] = L[i];
    g.lxh = L[0] / 2;
    g.Vh = (uint32_t)((uint64_t)L[0] * L[1] * L[2] * L[3] / 2);
    g.stride[0] = 1;
    g.stride[1] = (uint32_t)g.lxh;
    g.stride[2] = (uint32_t)g.lxh * L[1];
    g.stride[3] = (uint32_t)g.lxh * L[1] * L[2];
    g.dx = FastDiv((uint32_t)g.lxh);
    g.dy = FastDiv((uint32_t)L[1]);
    g.dz = FastDiv((uint32_t)L[2]);
    set_tiles(g, tile_slab_limit());
    return g;
}

struct SiteCoord {
    int c[4];  // full coordinates x, y, z, t
};

__host__ __device__ inline SiteCoord site_coords(const Geom &g, uint32_t s, int parity)
{
    SiteCoord sc;
    uint32_t r = g.dx.div(s);
    int xcb = (int)(s - r * (uint32_t)g.lxh);
    uint32_t r2 = g.dy.div(r);
    sc.c[1] = (int)(r - r2 * (uint32_t)g.L[1]);
    uint32_t r3 = g.dz.div(r2);
    sc.c[2] = (int)(r2 - r3 * (uint32_t)g.L[2]);
    sc.c[3] = (int)r3;
    sc.c[0] = 2 * xcb + ((sc.c[1] + sc.c[2] + sc.c[3] + parity) & 1);
    return sc;
}

// (c + h) mod L for |h| <= 3 and L >= 1, by at most three corrections.
__host__ __device__ inline int wrap_add(int c, int h, int L)
{
    int v = c + h;
    if (h > 0) {
        v = v >= L ? v - L : v;
        v = v >= L ? v - L : v;
        v = v >= L ? v - L : v;
    } else {
        v = v < 0 ? v + L : v;
        v = v < 0 ? v + L : v;
        v = v < 0 ? v + L : v;
    }
    return v;
}

// Checkerboard index of the neighbour of (s, coords) at hop h along mu.
template <int MU>
__host__ __device__ inline uint32_t nb_index(const Geom &g, uint32_t s, const SiteCoord &sc, int h)
{
    const int v = wrap_add(sc.c[MU], h, g.L[MU]);
    if (MU == 0) {
        // same y,z,t: only x>>1 changes
        return s - (uint32_t)(sc.c[0] >> 1) + (uint32_t)(v >> 1);
    } else {
        return (uint32_t)((int64_t)s + (int64_t)(v - sc.c[MU]) * (int64_t)g.stride[MU]);
    }
}

// ---------------------------------------------------------------- reductions
template <typename T> __device__ inline T wave_sum(T v)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1)
        v += __shfl_xor(v, off, kWave);
    return v;
}

// Sum over a workgroup of NT threads; result valid in thread 0.
// `lds` must hold NT/64 elements.
template <int NT, typename T> __device__ inline T block_sum(T v, T *lds)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    if (lane == 0)
        lds[wid] = v;
    __syncthreads();
    T r = 0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int w = 0; w < NT / kWave; ++w)
            r += lds[w];
    }
    return r;
}

// XCD-aware block remap: hardware dispatches workgroup b to XCD (b % 8).
// Remapping b -> (b % 8) * (nb / 8) + b / 8 gives every XCD one contiguous
// range of sites, so neighbour gathers along y/z/t stay in that XCD's L2.
// Valid when nb is a multiple of 8; otherwise identity.
__device__ inline uint32_t xcd_remap(uint32_t b, uint32_t nb)
{
    if ((nb & 7u) != 0)
        return b;
    return (b & 7u) * (nb >> 3) + (b >> 3);
}

// Site of thread tid of block b (of nb blocks of NT threads) for the stencil kernels.
// Workgroups are dealt to the 8 XCDs round robin (b % 8), each XCD with its own L2.
// When Vh is a multiple of 8*NT and the lattice splits into 8 tiles (make_geom: y ranges x
// z ranges x t ranges; 1 x 4 x 2 at 32^4, 2 x 4 x 1 at 48^3x96), XCD x owns one tile and walks
// it slab by slab ((x, y) within its y range fastest, then z, then t), so the +-1 and +-3
// neighbours along z and t that a site gathers were mostly just read by the same XCD.
// Otherwise each XCD gets a contiguous range of blocks (xcd_remap). Either way a bijection onto
// the sites (tile: v < Vh/8 = slab*zc*tc maps to distinct (p, zl, tl) per XCD).
template <int NT> __device__ inline uint32_t tile_site(const Geom &g, uint32_t b, uint32_t nb, uint32_t tid, bool &valid)
{
    if (g.tile && (g.Vh % (8u * NT)) == 0) {
        const uint32_t x = b & 7u, v = (b >> 3) * NT + tid;
        const uint32_t q = g.dslab.div(v);
        const uint32_t p = v - q * g.slab;
        const uint32_t tl = g.dzc.div(q);
        const uint32_t zl = q - tl * g.zc;
        const uint32_t xy = x & ((1u << g.lyp) - 1u), xz = (x >> g.lyp) & ((1u << g.lzp) - 1u);
        const uint32_t xt = x >> (g.lyp + g.lzp);
        valid = true;
        return p + g.slab * xy + g.plane * ((xz * g.zc + zl) + (uint32_t)g.L[2] * (xt * g.tc + tl));
    }
    const uint32_t s0 = xcd_remap(b, nb) * NT + tid;
    valid = s0 < g.Vh;
    return valid ? s0 : 0;
}

// x-rows of one (x, y) plane that each workgroup of nt threads covers in the tile order, when it
// covers whole rows of a single plane: the tile order is active, an XCD's slab is the whole
// plane (one y range), and nt is a multiple of the row length Lx/2 that divides the plane. The
// hopping terms then stage the workgroup's rows and their +-3 y-neighbour rows of every source
// in LDS (16-byte loads need an even row length). 0 when the workgroups do not have this shape.
__host__ inline int stage_rows(const Geom &g, int nt)
{
    if (!g.tile || g.Vh % (8u * (uint32_t)nt) != 0 || g.lyp != 0 || g.lxh <= 0 || g.lxh % 2 != 0 ||
        nt % g.lxh != 0 || g.plane % (uint32_t)nt != 0 || g.L[1] < 3)
        return 0;
    return nt / g.lxh;
}

}  // namespace kq
