// HIP kernels of the parallel-transport convolution preconditioner: the two
// forward layers (the second with an optional fused dot product for the
// flexible CG), the backward pass to the hidden layer, the weight gradients
// (deterministic two-stage FP64 reductions), and FP32 vector helpers.
#include <cstring>
#include <stdexcept>

#include "kq_ptnet.h"
#include "kq_solver.h"

namespace kq {

namespace {

constexpr int kNT = 256;  // threads per block of every kernel here

__device__ __forceinline__ float2 cmulf(float2 a, float2 b)
{
    return mk<float2>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ float2 unpack_w(uint32_t u)
{
    return mk<float2>(__uint_as_float(u << 16), __uint_as_float(u & 0xffff0000u));
}

// (re, im) -> bfloat16 pair in unpack_w's format, round to nearest even (finite values)
__device__ __forceinline__ uint32_t pack_bf16(float2 v)
{
    uint32_t a = __float_as_uint(v.x), b = __float_as_uint(v.y);
    a += 0x7fffu + ((a >> 16) & 1u);
    b += 0x7fffu + ((b >> 16) & 1u);
    return (a >> 16) | (b & 0xffff0000u);
}

// acc = L(s) x  or  acc = L(n)^+ x, for the tap's link format
template <class LK, bool NEG, typename R>
__device__ __forceinline__ void tap_mul(const LK &lk, int qo, int mu, uint32_t s, uint32_t n, const float2 x[3],
                                        R acc[6])
{
    float2 U[9];
#pragma unroll
    for (int k = 0; k < 6; ++k)
        acc[k] = 0;
    if (!NEG) {
        lk.load(qo, mu, s, U);
        umv_add(U, x, acc);
    } else {
        lk.load(1 - qo, mu, n, U);
        udmv_sub(U, x, acc);
#pragma unroll
        for (int k = 0; k < 6; ++k)
            acc[k] = -acc[k];
    }
}

// (T_t in)(s) for a one-channel field `in` on parity 1-qo
template <class LK, int MU, int H, bool NEG>
__device__ __forceinline__ void tap1(const Geom &g, int qo, const LK &lk, uint32_t s, const SiteCoord &sc,
                                     const float2 *in, float2 v[3])
{
    const uint32_t n = nb_index<MU>(g, s, sc, NEG ? -H : H);
    float2 x[3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
        x[c] = in[(size_t)c * g.Vh + n];
    float acc[6];
    tap_mul<LK, NEG>(lk, qo, MU, s, n, x, acc);
#pragma unroll
    for (int c = 0; c < 3; ++c)
        v[c] = mk<float2>(acc[2 * c], acc[2 * c + 1]);
}

// (T_t sum_ch wt[ch] in_ch)(s) for a kPtHidden-channel field
template <class LK, int MU, int H, bool NEG>
__device__ __forceinline__ void tap_comb(const Geom &g, int qo, const LK &lk, uint32_t s, const SiteCoord &sc,
                                         const float2 *in, const float2 wt[kPtHidden], float2 v[3])
{
    const uint32_t n = nb_index<MU>(g, s, sc, NEG ? -H : H);
    float2 x[3] = {mk<float2>(0.f, 0.f), mk<float2>(0.f, 0.f), mk<float2>(0.f, 0.f)};
#pragma unroll
    for (int ch = 0; ch < kPtHidden; ++ch) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float2 p = cmulf(wt[ch], in[((size_t)ch * 3 + c) * g.Vh + n]);
            x[c].x += p.x;
            x[c].y += p.y;
        }
    }
    float acc[6];
    tap_mul<LK, NEG>(lk, qo, MU, s, n, x, acc);
#pragma unroll
    for (int c = 0; c < 3; ++c)
        v[c] = mk<float2>(acc[2 * c], acc[2 * c + 1]);
}

// Layer 2 with bfloat16 channels, the four taps of direction MU (t = 4 MU + 0..3: hops +1,
// +3, -1, -3, as KQ_PT_FOR_TAPS orders them) added to acc in tap order. Every gather and link
// load of the four is issued before any of them is used: written tap by tap, the compiler
// waited for each 16-byte gather before issuing the next (81 of 102 loads followed by a full
// wait) and the kernel ran at a third of the HBM rate.
template <class LF, class LL, int MU>
__device__ __forceinline__ void taps16_dir(const Geom &g, int qo, const LF &fat, const LL &lng, uint32_t s,
                                           const SiteCoord &sc, const kq_v4u *in, const PtWeightsBF &w,
                                           float2 acc[3])
{
    const uint32_t n[4] = {nb_index<MU>(g, s, sc, 1), nb_index<MU>(g, s, sc, 3), nb_index<MU>(g, s, sc, -1),
                           nb_index<MU>(g, s, sc, -3)};
    kq_v4u u[4][3];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            u[t][c] = in[(size_t)c * g.Vh + n[t]];
    float2 U[4][9];
    fat.load(qo, MU, s, U[0]);
    lng.load(qo, MU, s, U[1]);
    fat.load(1 - qo, MU, n[2], U[2]);
    lng.load(1 - qo, MU, n[3], U[3]);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        float2 wt[kPtHidden];
#pragma unroll
        for (int c = 0; c < kPtHidden; ++c)
            wt[c] = unpack_w(w.w2[c][4 * MU + t]);
        float2 x[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float2 p0 = cmulf(wt[0], unpack_w(u[t][c].x)), p1 = cmulf(wt[1], unpack_w(u[t][c].y));
            const float2 p2 = cmulf(wt[2], unpack_w(u[t][c].z)), p3 = cmulf(wt[3], unpack_w(u[t][c].w));
            x[c] = mk<float2>((p0.x + p1.x) + (p2.x + p3.x), (p0.y + p1.y) + (p2.y + p3.y));
        }
        float v[6];
#pragma unroll
        for (int k = 0; k < 6; ++k)
            v[k] = 0;
        if (t < 2) {
            umv_add(U[t], x, v);
        } else {
            udmv_sub(U[t], x, v);
#pragma unroll
            for (int k = 0; k < 6; ++k)
                v[k] = -v[k];
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            acc[k].x += v[2 * k];
            acc[k].y += v[2 * k + 1];
        }
    }
}

// the 16 taps as (t, mu, hop, negative)
#define KQ_PT_FOR_TAPS(X)                                                                                       \
    X(0, 0, 1, false) X(1, 0, 3, false) X(2, 0, 1, true) X(3, 0, 3, true) X(4, 1, 1, false) X(5, 1, 3, false)   \
        X(6, 1, 1, true) X(7, 1, 3, true) X(8, 2, 1, false) X(9, 2, 3, false) X(10, 2, 1, true)                \
            X(11, 2, 3, true) X(12, 3, 1, false) X(13, 3, 3, false) X(14, 3, 1, true) X(15, 3, 3, true)

struct PtLayerArgs {
    Geom g;
    int qo;              // output parity
    const float2 *in;    // input field(s) on parity 1-qo, channel stride 3*Vh
    float2 *out;         // output field(s) on parity qo, channel stride 3*Vh (may be null with out16)
    kq_v4u *out16;       // layer 1: the outputs also as bfloat16 (pt_forward's h16 layout); may be null
    const kq_v4u *in16;  // layer 2: read the channels from here (bfloat16) instead of in; may be null
    const float2 *skip;  // layer 2: w0 * skip(x) added (parity qo); may be null
    const float2 *dotq;  // layer 2: per-block partials of <dotq, out> (may be null)
    double *part;        // [2*nblk]
    const int *stop;     // do nothing while *stop != 0 (may be null)
    PtWeightsBF w;       // layer-1 form uses w1; layer-2 form uses w2 and w0
};

// layer-1 form: out_c(x) = sum_t w1[c][t] (T_t in)(x), one input channel, kPtHidden outputs
template <class LF, class LL>
__global__ void __launch_bounds__(kNT) k_pt_layer1(const PtLayerArgs a, LF fat, LL lng)
{
    if (a.stop && *a.stop)
        return;
    bool valid;
    const uint32_t s = tile_site<kNT>(a.g, blockIdx.x, gridDim.x, threadIdx.x, valid);
    if (!valid)
        return;
    const SiteCoord sc = site_coords(a.g, s, a.qo);
    float2 acc[kPtHidden][3];
#pragma unroll
    for (int c = 0; c < kPtHidden; ++c)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            acc[c][k] = mk<float2>(0.f, 0.f);
#define KQ_TAP1(T, MU, H, NEG)                                                    \
    {                                                                             \
        float2 v[3];                                                              \
        if (H == 1)                                                               \
            tap1<LF, MU, H, NEG>(a.g, a.qo, fat, s, sc, a.in, v);                 \
        else                                                                      \
            tap1<LL, MU, H, NEG>(a.g, a.qo, lng, s, sc, a.in, v);                 \
        _Pragma("unroll") for (int c = 0; c < kPtHidden; ++c)                     \
        {                                                                         \
            const float2 w = unpack_w(a.w.w1[c][T]);                              \
            _Pragma("unroll") for (int k = 0; k < 3; ++k)                         \
            {                                                                     \
                const float2 p = cmulf(w, v[k]);                                  \
                acc[c][k].x += p.x;                                               \
                acc[c][k].y += p.y;                                               \
            }                                                                     \
        }                                                                         \
    }
    KQ_PT_FOR_TAPS(KQ_TAP1)
#undef KQ_TAP1
    if (a.out) {
#pragma unroll
        for (int c = 0; c < kPtHidden; ++c)
#pragma unroll
            for (int k = 0; k < 3; ++k)
                a.out[((size_t)c * 3 + k) * a.g.Vh + s] = acc[c][k];
    }
    if (a.out16) {
        static_assert(kPtHidden == 4, "the bfloat16 hidden layout holds 4 channels");
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            kq_v4u u;
            u.x = pack_bf16(acc[0][k]);
            u.y = pack_bf16(acc[1][k]);
            u.z = pack_bf16(acc[2][k]);
            u.w = pack_bf16(acc[3][k]);
            a.out16[(size_t)k * a.g.Vh + s] = u;
        }
    }
}

// layer-2 form: out(x) = w0 skip(x) + sum_t T_t( sum_c w2[c][t] in_c )(x); H16: channels from in16
template <class LF, class LL, bool H16>
__global__ void __launch_bounds__(kNT) __attribute__((amdgpu_waves_per_eu(1, 4)))
k_pt_layer2(const PtLayerArgs a, LF fat, LL lng)
{
    if (a.stop && *a.stop)
        return;
    bool valid;
    const uint32_t s = tile_site<kNT>(a.g, blockIdx.x, gridDim.x, threadIdx.x, valid);
    const SiteCoord sc = site_coords(a.g, s, a.qo);
    float2 acc[3];
    const float2 w0 = unpack_w(a.w.w0);
#pragma unroll
    for (int k = 0; k < 3; ++k)
        acc[k] = a.skip ? cmulf(w0, a.skip[(size_t)k * a.g.Vh + s]) : mk<float2>(0.f, 0.f);
    if constexpr (H16) {
        taps16_dir<LF, LL, 0>(a.g, a.qo, fat, lng, s, sc, a.in16, a.w, acc);
        taps16_dir<LF, LL, 1>(a.g, a.qo, fat, lng, s, sc, a.in16, a.w, acc);
        taps16_dir<LF, LL, 2>(a.g, a.qo, fat, lng, s, sc, a.in16, a.w, acc);
        taps16_dir<LF, LL, 3>(a.g, a.qo, fat, lng, s, sc, a.in16, a.w, acc);
    } else {
#define KQ_TAP2(T, MU, H, NEG)                                                     \
    {                                                                              \
        float2 wt[kPtHidden], v[3];                                                \
        _Pragma("unroll") for (int c = 0; c < kPtHidden; ++c) wt[c] = unpack_w(a.w.w2[c][T]); \
        if (H == 1)                                                                \
            tap_comb<LF, MU, H, NEG>(a.g, a.qo, fat, s, sc, a.in, wt, v);          \
        else                                                                       \
            tap_comb<LL, MU, H, NEG>(a.g, a.qo, lng, s, sc, a.in, wt, v);          \
        _Pragma("unroll") for (int k = 0; k < 3; ++k)                              \
        {                                                                          \
            acc[k].x += v[k].x;                                                    \
            acc[k].y += v[k].y;                                                    \
        }                                                                          \
    }
        KQ_PT_FOR_TAPS(KQ_TAP2)
#undef KQ_TAP2
    }
    if (valid) {
#pragma unroll
        for (int k = 0; k < 3; ++k)
            a.out[(size_t)k * a.g.Vh + s] = acc[k];
    }
    if (a.dotq) {
        __shared__ double lre[kNT / kWave], lim[kNT / kWave];
        double re = 0.0, im = 0.0;
        if (valid) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float2 d = a.dotq[(size_t)k * a.g.Vh + s];
                re += (double)d.x * acc[k].x + (double)d.y * acc[k].y;
                im += (double)d.x * acc[k].y - (double)d.y * acc[k].x;
            }
        }
        re = wave_sum(re);
        im = wave_sum(im);
        if ((threadIdx.x & (kWave - 1)) == 0) {
            lre[threadIdx.x / kWave] = re;
            lim[threadIdx.x / kWave] = im;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double sr = 0.0, si = 0.0;
            for (int w = 0; w < kNT / kWave; ++w) {
                sr += lre[w];
                si += lim[w];
            }
            a.part[2 * blockIdx.x] = sr;
            a.part[2 * blockIdx.x + 1] = si;
        }
    }
}

// P[c][t] = sum over sites y of parity qo of <(T_t in)(y), b_c(y)>;
// part[blk][c*16 + t] (complex, FP64)
template <class LF, class LL>
__global__ void __launch_bounds__(kNT) k_pt_wgrad(Geom g, int qo, LF fat, LL lng, const float2 *in,
                                                  const float2 *b, double2 *part)
{
    __shared__ double lds[kNT / kWave][kPtHidden * kPtTaps * 2];
    bool valid;
    const uint32_t s = tile_site<kNT>(g, blockIdx.x, gridDim.x, threadIdx.x, valid);
    const SiteCoord sc = site_coords(g, s, qo);
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    float2 bb[kPtHidden][3];
#pragma unroll
    for (int c = 0; c < kPtHidden; ++c)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            bb[c][k] = valid ? b[((size_t)c * 3 + k) * g.Vh + s] : mk<float2>(0.f, 0.f);
#define KQ_TAPG(T, MU, H, NEG)                                                    \
    {                                                                             \
        float2 v[3];                                                              \
        if (H == 1)                                                               \
            tap1<LF, MU, H, NEG>(g, qo, fat, s, sc, in, v);                       \
        else                                                                      \
            tap1<LL, MU, H, NEG>(g, qo, lng, s, sc, in, v);                       \
        _Pragma("unroll") for (int c = 0; c < kPtHidden; ++c)                     \
        {                                                                         \
            double re = 0.0, im = 0.0;                                            \
            _Pragma("unroll") for (int k = 0; k < 3; ++k)                         \
            {                                                                     \
                re += (double)v[k].x * bb[c][k].x + (double)v[k].y * bb[c][k].y;  \
                im += (double)v[k].x * bb[c][k].y - (double)v[k].y * bb[c][k].x;  \
            }                                                                     \
            re = wave_sum(re);                                                    \
            im = wave_sum(im);                                                    \
            if (lane == 0) {                                                      \
                lds[wid][2 * (c * kPtTaps + T)] = re;                             \
                lds[wid][2 * (c * kPtTaps + T) + 1] = im;                         \
            }                                                                     \
        }                                                                         \
    }
    KQ_PT_FOR_TAPS(KQ_TAPG)
#undef KQ_TAPG
    __syncthreads();
    if (threadIdx.x < kPtHidden * kPtTaps) {
        double re = 0.0, im = 0.0;
        for (int w = 0; w < kNT / kWave; ++w) {
            re += lds[w][2 * threadIdx.x];
            im += lds[w][2 * threadIdx.x + 1];
        }
        part[(size_t)blockIdx.x * kPtHidden * kPtTaps + threadIdx.x] = mk<double2>(re, im);
    }
}

// out[j] = sum over blocks b of part[b*stride + j], j = blockIdx.x; fixed summation order
__global__ void __launch_bounds__(kNT) k_reduce_cols(const double2 *part, uint32_t nblk, uint32_t stride,
                                                     double2 *out)
{
    const int j = blockIdx.x;
    __shared__ double lre[kNT], lim[kNT];
    double re = 0.0, im = 0.0;
    for (uint32_t b = threadIdx.x; b < nblk; b += kNT) {
        const double2 v = part[(size_t)b * stride + j];
        re += v.x;
        im += v.y;
    }
    lre[threadIdx.x] = re;
    lim[threadIdx.x] = im;
    __syncthreads();
    for (int w = kNT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            lre[threadIdx.x] += lre[threadIdx.x + w];
            lim[threadIdx.x] += lim[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        out[j] = mk<double2>(lre[0], lim[0]);
}

__global__ void k_vf_axpy(float2 *y, const float2 *x, float2 a, size_t n)
{
    const size_t i = (size_t)blockIdx.x * kNT + threadIdx.x;
    if (i < n) {
        const float2 p = cmulf(a, x[i]);
        y[i] = mk<float2>(y[i].x + p.x, y[i].y + p.y);
    }
}

__global__ void k_vf_lincomb(float2 *y, const float2 *x1, float2 a1, const float2 *x2, float2 a2, size_t n)
{
    const size_t i = (size_t)blockIdx.x * kNT + threadIdx.x;
    if (i < n) {
        const float2 p = cmulf(a1, x1[i]), q = cmulf(a2, x2[i]);
        y[i] = mk<float2>(p.x + q.x, p.y + q.y);
    }
}

struct DotSet {
    const float2 *x[8];
    const float2 *y[8];
};

// part[k*nblk + blk] = block partial of <x_k, y_k>
__global__ void __launch_bounds__(kNT) k_vf_dots(DotSet d, int nd, size_t n, double2 *part)
{
    __shared__ double lre[kNT / kWave], lim[kNT / kWave];
    const size_t i = (size_t)blockIdx.x * kNT + threadIdx.x;
    for (int k = 0; k < nd; ++k) {
        double re = 0.0, im = 0.0;
        if (i < n) {
            const float2 a = d.x[k][i], b = d.y[k][i];
            re = (double)a.x * b.x + (double)a.y * b.y;
            im = (double)a.x * b.y - (double)a.y * b.x;
        }
        re = wave_sum(re);
        im = wave_sum(im);
        if ((threadIdx.x & (kWave - 1)) == 0) {
            lre[threadIdx.x / kWave] = re;
            lim[threadIdx.x / kWave] = im;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double sr = 0.0, si = 0.0;
            for (int w = 0; w < kNT / kWave; ++w) {
                sr += lre[w];
                si += lim[w];
            }
            part[(size_t)k * gridDim.x + blockIdx.x] = mk<double2>(sr, si);
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kNT) k_vd_sumsq(const double *x, size_t n, double *part)
{
    __shared__ double l[kNT / kWave];
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * kNT + threadIdx.x; i < n; i += (size_t)gridDim.x * kNT)
        acc += x[i] * x[i];
    acc = wave_sum(acc);
    if ((threadIdx.x & (kWave - 1)) == 0)
        l[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kNT / kWave; ++w)
            s += l[w];
        part[blockIdx.x] = s;
    }
}

__global__ void __launch_bounds__(kNT) k_unpack_hidden(const kq_v4u *h16, float2 *h, uint32_t Vh)
{
    const size_t i = (size_t)blockIdx.x * kNT + threadIdx.x;  // colour-major entry k*Vh + s
    if (i >= (size_t)3 * Vh)
        return;
    const kq_v4u u = h16[i];
    const size_t n3 = (size_t)3 * Vh;
    h[i] = unpack_w(u.x);
    h[n3 + i] = unpack_w(u.y);
    h[2 * n3 + i] = unpack_w(u.z);
    h[3 * n3 + i] = unpack_w(u.w);
}

inline unsigned grid_of(size_t n) { return (unsigned)((n + kNT - 1) / kNT); }

}  // namespace

void pt_unpack_hidden(const kq_v4u *h16, float2 *h, uint32_t Vh, hipStream_t st)
{
    if (Vh == 0)
        return;
    hipLaunchKernelGGL(k_unpack_hidden, dim3(grid_of((size_t)3 * Vh)), dim3(kNT), 0, st, h16, h, Vh);
    KQ_HIP_CHECK(hipGetLastError());
}

uint32_t pt_layer_nblk(uint32_t Vh) { return (Vh + kNT - 1) / kNT; }

template <class LF, class LL>
void pt_forward(const Geom &g, int q, LF fat, LL lng, const PtWeightsBF &w, const float2 *r, float2 *h, float2 *z,
                hipStream_t st, const float2 *dotq, double *part, const int *stop, kq_v4u *h16)
{
    if (g.Vh == 0)
        return;
    const unsigned nb = pt_layer_nblk(g.Vh);
    PtLayerArgs a{};
    a.g = g;
    a.w = w;
    a.stop = stop;
    a.qo = 1 - q;
    a.in = r;
    a.out = h;
    a.out16 = h16;
    hipLaunchKernelGGL((k_pt_layer1<LF, LL>), dim3(nb), dim3(kNT), 0, st, a, fat, lng);
    KQ_HIP_CHECK(hipGetLastError());
    a.qo = q;
    a.in = h;
    a.in16 = h16;
    a.out16 = nullptr;
    a.out = z;
    a.skip = r;
    a.dotq = dotq;
    a.part = part;
    if (h16)
        hipLaunchKernelGGL((k_pt_layer2<LF, LL, true>), dim3(nb), dim3(kNT), 0, st, a, fat, lng);
    else
        hipLaunchKernelGGL((k_pt_layer2<LF, LL, false>), dim3(nb), dim3(kNT), 0, st, a, fat, lng);
    KQ_HIP_CHECK(hipGetLastError());
}

template <class LF, class LL>
void pt_backward_hidden(const Geom &g, int q, LF fat, LL lng, const PtWeightsBF &wb, const float2 *gz, float2 *gh,
                        hipStream_t st)
{
    if (g.Vh == 0)
        return;
    PtLayerArgs a{};
    a.g = g;
    a.w = wb;
    a.qo = 1 - q;
    a.in = gz;
    a.out = gh;
    hipLaunchKernelGGL((k_pt_layer1<LF, LL>), dim3(pt_layer_nblk(g.Vh)), dim3(kNT), 0, st, a, fat, lng);
    KQ_HIP_CHECK(hipGetLastError());
}

template <class LF, class LL>
static void wgrad_pass(const Geom &g, int qo, LF fat, LL lng, const float2 *in, const float2 *b, double2 *res,
                       hipStream_t st)
{
    const unsigned nb = pt_layer_nblk(g.Vh);
    constexpr int J = kPtHidden * kPtTaps;
    DevBuf part(sizeof(double2) * nb * J), out(sizeof(double2) * J);
    hipLaunchKernelGGL((k_pt_wgrad<LF, LL>), dim3(nb), dim3(kNT), 0, st, g, qo, fat, lng, in, b, part.as<double2>());
    KQ_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_cols, dim3(J), dim3(kNT), 0, st, part.as<double2>(), nb, (uint32_t)J,
                       out.as<double2>());
    KQ_HIP_CHECK(hipGetLastError());
    KQ_HIP_CHECK(hipMemcpyAsync(res, out.as<void>(), sizeof(double2) * J, hipMemcpyDeviceToHost, st));
    KQ_HIP_CHECK(hipStreamSynchronize(st));
}

template <class LF, class LL>
void pt_weight_grads(const Geom &g, int q, LF fat, LL lng, const float2 *r, const float2 *h, const float2 *gz,
                     const float2 *gh, double *out, hipStream_t st)
{
    constexpr int J = kPtHidden * kPtTaps;
    double2 p1[J], p2[J];
    // gW1[c][t] = sum_y <T_t r, gh_c>, y on parity 1-q
    wgrad_pass<LF, LL>(g, 1 - q, fat, lng, r, gh, p1, st);
    // gW2[c][t] = sum_y <h_c, T_{flip t} g> = conj(sum_y <T_{flip t} g, h_c>)
    wgrad_pass<LF, LL>(g, 1 - q, fat, lng, gz, h, p2, st);
    for (int c = 0; c < kPtHidden; ++c)
        for (int t = 0; t < kPtTaps; ++t) {
            const int j1 = c * kPtTaps + t, j2 = J + c * kPtTaps + t;
            out[2 * j1] = p1[j1].x;
            out[2 * j1 + 1] = p1[j1].y;
            const double2 b = p2[c * kPtTaps + pt_flip_tap(t)];
            out[2 * j2] = b.x;
            out[2 * j2 + 1] = -b.y;
        }
    // gW0 = sum_x <r, g>
    double2 d0;
    const float2 *xs[1] = {r};
    const float2 *ys[1] = {gz};
    vf_dots(xs, ys, 1, (size_t)3 * g.Vh, &d0, st);
    out[4 * J] = d0.x;
    out[4 * J + 1] = d0.y;
}

// ---------------------------------------------------------------- weights
float bf16_round(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u)
        return f;  // inf / nan unchanged
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest, ties to even
    u &= 0xffff0000u;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

uint32_t bf16_pack(float2 w)
{
    const float re = bf16_round(w.x), im = bf16_round(w.y);
    uint32_t ur, ui;
    std::memcpy(&ur, &re, 4);
    std::memcpy(&ui, &im, 4);
    return (ur >> 16) | (ui & 0xffff0000u);
}

float2 bf16_unpack(uint32_t u)
{
    const uint32_t ur = u << 16, ui = u & 0xffff0000u;
    float re, im;
    std::memcpy(&re, &ur, 4);
    std::memcpy(&im, &ui, 4);
    return mk<float2>(re, im);
}

static float2 scaled(float2 w, double s) { return mk<float2>((float)(w.x * s), (float)(w.y * s)); }

PtWeightsBF pt_pack(const PtWeights &w, const PtNorms &nu)
{
    PtWeightsBF o{};
    for (int c = 0; c < kPtHidden; ++c)
        for (int t = 0; t < kPtTaps; ++t) {
            const double s = 1.0 / nu.nu[pt_tap_hop(t) == 3 ? 1 : 0];
            o.w1[c][t] = bf16_pack(scaled(w.w1[c][t], s));
            o.w2[c][t] = bf16_pack(scaled(w.w2[c][t], s));
        }
    o.w0 = bf16_pack(w.w0);
    return o;
}

PtWeightsBF pt_pack_backward(const PtWeights &w, const PtNorms &nu)
{
    const PtWeights e = pt_effective(w, nu);
    PtWeightsBF o{};
    for (int c = 0; c < kPtHidden; ++c)
        for (int t = 0; t < kPtTaps; ++t) {
            const float2 v = e.w2[c][pt_flip_tap(t)];
            o.w1[c][t] = bf16_pack(mk<float2>(v.x, -v.y));  // exact: v is already bf16
        }
    return o;
}

PtWeights pt_effective(const PtWeights &w, const PtNorms &nu)
{
    const PtWeightsBF p = pt_pack(w, nu);
    PtWeights o{};
    for (int c = 0; c < kPtHidden; ++c)
        for (int t = 0; t < kPtTaps; ++t) {
            o.w1[c][t] = bf16_unpack(p.w1[c][t]);
            o.w2[c][t] = bf16_unpack(p.w2[c][t]);
        }
    o.w0 = bf16_unpack(p.w0);
    return o;
}

// ---------------------------------------------------------------- vector helpers
void vf_axpy(float2 *y, const float2 *x, float2 a, size_t n, hipStream_t st)
{
    if (n == 0)
        return;
    hipLaunchKernelGGL(k_vf_axpy, dim3(grid_of(n)), dim3(kNT), 0, st, y, x, a, n);
    KQ_HIP_CHECK(hipGetLastError());
}

void vf_lincomb(float2 *y, const float2 *x1, float2 a1, const float2 *x2, float2 a2, size_t n, hipStream_t st)
{
    if (n == 0)
        return;
    hipLaunchKernelGGL(k_vf_lincomb, dim3(grid_of(n)), dim3(kNT), 0, st, y, x1, a1, x2, a2, n);
    KQ_HIP_CHECK(hipGetLastError());
}

void vf_dots(const float2 *const *xs, const float2 *const *ys, int ndot, size_t n, double2 *out, hipStream_t st)
{
    if (ndot <= 0 || ndot > 8)
        throw std::invalid_argument("vf_dots: 1..8 dot products per launch");
    if (n == 0) {
        for (int k = 0; k < ndot; ++k)
            out[k] = mk<double2>(0.0, 0.0);
        return;
    }
    DotSet d{};
    for (int k = 0; k < ndot; ++k) {
        d.x[k] = xs[k];
        d.y[k] = ys[k];
    }
    const unsigned nb = grid_of(n);
    DevBuf part(sizeof(double2) * ndot * nb), res(sizeof(double2) * ndot);
    hipLaunchKernelGGL(k_vf_dots, dim3(nb), dim3(kNT), 0, st, d, ndot, n, part.as<double2>());
    KQ_HIP_CHECK(hipGetLastError());
    for (int k = 0; k < ndot; ++k) {
        hipLaunchKernelGGL(k_reduce_cols, dim3(1), dim3(kNT), 0, st, part.as<double2>() + (size_t)k * nb, nb, 1u,
                           res.as<double2>() + k);
        KQ_HIP_CHECK(hipGetLastError());
    }
    KQ_HIP_CHECK(hipMemcpyAsync(out, res.as<void>(), sizeof(double2) * ndot, hipMemcpyDeviceToHost, st));
    KQ_HIP_CHECK(hipStreamSynchronize(st));
}

double vd_sumsq(const double *x, size_t n, hipStream_t st)
{
    if (n == 0)
        return 0.0;
    const unsigned nb = (unsigned)std::min<size_t>(1024, grid_of(n));
    DevBuf part(sizeof(double) * nb);
    hipLaunchKernelGGL(k_vd_sumsq, dim3(nb), dim3(kNT), 0, st, x, n, part.as<double>());
    KQ_HIP_CHECK(hipGetLastError());
    std::vector<double> h(nb);
    KQ_HIP_CHECK(hipMemcpyAsync(h.data(), part.as<void>(), sizeof(double) * nb, hipMemcpyDeviceToHost, st));
    KQ_HIP_CHECK(hipStreamSynchronize(st));
    double s = 0.0;
    for (double v : h)
        s += v;
    return s;
}

#define KQ_PT_INST(LF, LL)                                                                                          \
    template void pt_forward<LF, LL>(const Geom &, int, LF, LL, const PtWeightsBF &, const float2 *, float2 *,       \
                                     float2 *, hipStream_t, const float2 *, double *, const int *, kq_v4u *);       \
    template void pt_backward_hidden<LF, LL>(const Geom &, int, LF, LL, const PtWeightsBF &, const float2 *,         \
                                             float2 *, hipStream_t);                                                \
    template void pt_weight_grads<LF, LL>(const Geom &, int, LF, LL, const float2 *, const float2 *,                 \
                                          const float2 *, const float2 *, double *, hipStream_t);
KQ_PT_INST(LinkF, LinkF)
KQ_PT_INST(LinkF, LinkF12)
KQ_PT_INST(LinkH, LinkH12)
KQ_PT_INST(LinkH, LinkH)
KQ_PT_INST(LinkF, LinkH12)
KQ_PT_INST(LinkF, LinkH)
KQ_PT_INST(LinkF, LinkF12C)
KQ_PT_INST(LinkH, LinkH12C)
KQ_PT_INST(LinkF, LinkH12C)

}  // namespace kq
