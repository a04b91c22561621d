// Layout and precision conversion kernels (run once per gauge load and per
// solve for the source/solutions; not on the iteration path).
#include "kq_layout.h"

#include <algorithm>

namespace kq {

template <typename HR, typename LR>
__global__ void k_gauge_host_to_soa(const HR *raw, size_t i0, size_t n, c2_t<LR> *dst, uint32_t Vh, double scale)
{
    // one thread per (site, mu, k) of the chunk; consecutive threads read consecutive complex entries
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * 36)
        return;
    const size_t site = t / 36;
    const int rem = (int)(t - site * 36);
    const int mu = rem / 9, k = rem - mu * 9;
    const size_t i = i0 + site;
    const uint32_t q = (uint32_t)(i / Vh);
    const uint32_t s = (uint32_t)(i - (size_t)q * Vh);
    const HR re = raw[2 * t], im = raw[2 * t + 1];
    c2_t<LR> v;
    v.x = (LR)((double)re * scale);
    v.y = (LR)((double)im * scale);
    dst[((size_t)(q * 4 + mu) * 9 + k) * Vh + s] = v;
}

template <typename HR, typename LR>
void gauge_host_to_soa(const HR *raw, size_t i0, size_t n, c2_t<LR> *dst, uint32_t Vh, double scale,
                       hipStream_t st)
{
    const size_t nt = n * 36;
    if (nt == 0)
        return;
    hipLaunchKernelGGL((k_gauge_host_to_soa<HR, LR>), dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, st, raw, i0,
                       n, dst, Vh, scale);
    KQ_HIP_CHECK(hipGetLastError());
}

// one thread per link (q, mu, s)
__global__ void k_canon_to_linkD(const double2 *canon, char *dst, uint32_t Vh, uint32_t nb64)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)8 * Vh)
        return;
    const uint32_t qm = (uint32_t)(t / Vh), s = (uint32_t)(t - (size_t)qm * Vh);
    double2 *b = reinterpret_cast<double2 *>(dst + link_block_offset(qm / 4, qm % 4, s, nb64, LinkD::kBlock)) + (s & 63);
    for (int k = 0; k < 9; ++k)
        b[64 * k] = canon[((size_t)qm * 9 + k) * Vh + s];
}

void canon_to_linkD(const double2 *canon, char *dst, uint32_t Vh, hipStream_t st)
{
    const size_t n = (size_t)8 * Vh;
    if (n == 0)
        return;
    hipLaunchKernelGGL(k_canon_to_linkD, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, canon, dst, Vh,
                       nb64_of(Vh));
    KQ_HIP_CHECK(hipGetLastError());
}

template <typename CR>
__global__ void k_canon_to_linkF(const c2_t<CR> *canon, char *dst, uint32_t Vh, uint32_t nb64)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)8 * Vh)
        return;
    const uint32_t qm = (uint32_t)(t / Vh), s = (uint32_t)(t - (size_t)qm * Vh);
    char *b = dst + link_block_offset(qm / 4, qm % 4, s, nb64, LinkF::kBlock);
    const uint32_t l = s & 63;
    float v[18];
    for (int k = 0; k < 9; ++k) {
        const c2_t<CR> c = canon[((size_t)qm * 9 + k) * Vh + s];
        v[2 * k] = (float)c.x;
        v[2 * k + 1] = (float)c.y;
    }
    for (int c = 0; c < 4; ++c)
        reinterpret_cast<float4 *>(b + 1024 * c)[l] = make_float4(v[4 * c], v[4 * c + 1], v[4 * c + 2], v[4 * c + 3]);
    reinterpret_cast<float2 *>(b + 4096)[l] = make_float2(v[16], v[17]);
}

template <typename CR> void canon_to_linkF(const c2_t<CR> *canon, char *dst, uint32_t Vh, hipStream_t st)
{
    const size_t n = (size_t)8 * Vh;
    if (n == 0)
        return;
    hipLaunchKernelGGL((k_canon_to_linkF<CR>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, canon, dst, Vh,
                       nb64_of(Vh));
    KQ_HIP_CHECK(hipGetLastError());
}

// two reals in [-1, 1] -> 16-bit fixed point q = round(32767 x), real part in the low half
__device__ inline unsigned int pack_q2(double a, double b)
{
    const int qa = (int)rint(fmin(fmax(a, -1.0), 1.0) * 32767.0);
    const int qb = (int)rint(fmin(fmax(b, -1.0), 1.0) * 32767.0);
    return ((unsigned int)(unsigned short)(short)qa) | ((unsigned int)(unsigned short)(short)qb << 16);
}

__global__ void k_canon_to_linkH(const double2 *canon, char *dst, uint32_t Vh, uint32_t nb64)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)8 * Vh)
        return;
    const uint32_t qm = (uint32_t)(t / Vh), s = (uint32_t)(t - (size_t)qm * Vh);
    double2 m[9];
    double amax = 0.0;
    for (int k = 0; k < 9; ++k) {
        m[k] = canon[((size_t)qm * 9 + k) * Vh + s];
        amax = fmax(amax, fmax(fabs(m[k].x), fabs(m[k].y)));
    }
    const double a = amax > 0.0 ? amax : 1.0;
    unsigned int h[9];
    for (int k = 0; k < 9; ++k)
        h[k] = pack_q2(m[k].x / a, m[k].y / a);
    char *b = dst + link_block_offset(qm / 4, qm % 4, s, nb64, LinkH::kBlock);
    const uint32_t l = s & 63;
    reinterpret_cast<uint4 *>(b)[l] = make_uint4(h[0], h[1], h[2], h[3]);
    reinterpret_cast<uint4 *>(b + 1024)[l] = make_uint4(h[4], h[5], h[6], h[7]);
    reinterpret_cast<uint2 *>(b + 2048)[l] = make_uint2(h[8], __float_as_uint((float)(a / 32767.0)));
}

void canon_to_linkH(const double2 *canon, char *dst, uint32_t Vh, hipStream_t st)
{
    const size_t n = (size_t)8 * Vh;
    if (n == 0)
        return;
    hipLaunchKernelGGL(k_canon_to_linkH, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, canon, dst, Vh,
                       nb64_of(Vh));
    KQ_HIP_CHECK(hipGetLastError());
}

// det of a 3x3 complex matrix m[row*3+col]
__device__ inline double2 det3(const double2 m[9])
{
    auto mul = [](double2 a, double2 b) { return mk<double2>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); };
    auto sub = [](double2 a, double2 b) { return mk<double2>(a.x - b.x, a.y - b.y); };
    auto add = [](double2 a, double2 b) { return mk<double2>(a.x + b.x, a.y + b.y); };
    const double2 c0 = sub(mul(m[4], m[8]), mul(m[5], m[7]));
    const double2 c1 = sub(mul(m[5], m[6]), mul(m[3], m[8]));
    const double2 c2 = sub(mul(m[3], m[7]), mul(m[4], m[6]));
    return add(add(mul(m[0], c0), mul(m[1], c1)), mul(m[2], c2));
}

// V = U / b with b = cbrt(Re det U), checked to be unitary; false if U is not b * SU(3).
// C2: double2 or float2 canonical links (formed in double either way)
template <typename C2>
__device__ inline bool su3_factor(const C2 *canon, uint32_t qm, uint32_t s, uint32_t Vh, double2 v[9], double &bsc)
{
    double2 m[9];
    for (int k = 0; k < 9; ++k) {
        const C2 c = canon[((size_t)qm * 9 + k) * Vh + s];
        m[k] = mk<double2>((double)c.x, (double)c.y);
    }
    const double2 d = det3(m);
    // b^3 = det(b V) must be real; b = cbrt(Re det)
    bsc = cbrt(d.x);
    bool ok = fabs(d.y) <= 1e-6 * fabs(d.x) && bsc != 0.0;
    for (int k = 0; k < 9; ++k)
        v[k] = ok ? mk<double2>(m[k].x / bsc, m[k].y / bsc) : mk<double2>(0.0, 0.0);
    if (ok) {
        // unitarity of V: rows orthonormal
        for (int r1 = 0; r1 < 3 && ok; ++r1)
            for (int r2 = r1; r2 < 3 && ok; ++r2) {
                double re = 0.0, im = 0.0;
                for (int c = 0; c < 3; ++c) {
                    const double2 a = v[3 * r1 + c], bb = v[3 * r2 + c];
                    re += a.x * bb.x + a.y * bb.y;
                    im += a.y * bb.x - a.x * bb.y;
                }
                const double want = (r1 == r2) ? 1.0 : 0.0;
                if (fabs(re - want) > 1e-6 || fabs(im) > 1e-6)
                    ok = false;
            }
    }
    return ok;
}

__global__ void k_canon_to_linkH12(const double2 *canon, char *dst, uint32_t Vh, uint32_t nb64, int *fail)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)8 * Vh)
        return;
    const uint32_t qm = (uint32_t)(t / Vh), s = (uint32_t)(t - (size_t)qm * Vh);
    double2 v[9];
    double bsc;
    if (!su3_factor(canon, qm, s, Vh, v, bsc))
        atomicAdd(fail, 1);
    unsigned int h[6];
    for (int k = 0; k < 6; ++k)
        h[k] = pack_q2(v[k].x, v[k].y);
    char *b = dst + link_block_offset(qm / 4, qm % 4, s, nb64, LinkH12::kBlock);
    const uint32_t l = s & 63;
    reinterpret_cast<uint4 *>(b)[l] = make_uint4(h[0], h[1], h[2], h[3]);
    // complex 4, 5 and the scale as one 12-byte record
    unsigned int *rec = reinterpret_cast<unsigned int *>(b + 1024 + 12 * l);
    rec[0] = h[4];
    rec[1] = h[5];
    rec[2] = __float_as_uint((float)bsc);
}

// U = w V with w = principal cube root of det U, V = U / w; true when the rebuilt w (V0, V1,
// conj(V0 x V1)) matches U to tol * max|U_ij| (U a complex multiple of an SU(3) matrix)
template <typename C2>
__device__ inline bool su3_factor_c(const C2 *canon, uint32_t qm, uint32_t s, uint32_t Vh, double2 v[6], double2 &w,
                                    double tol)
{
    double2 m[9];
    double amax = 0.0;
    for (int k = 0; k < 9; ++k) {
        const C2 c = canon[((size_t)qm * 9 + k) * Vh + s];
        m[k] = mk<double2>((double)c.x, (double)c.y);
        amax = fmax(amax, fmax(fabs(m[k].x), fabs(m[k].y)));
    }
    const double2 d = det3(m);
    const double r = cbrt(hypot(d.x, d.y)), ph = atan2(d.y, d.x) / 3.0;
    w = mk<double2>(r * cos(ph), r * sin(ph));
    const double inv = r > 0.0 ? 1.0 / (r * r) : 0.0;
    for (int k = 0; k < 6; ++k)
        v[k] = mk<double2>((m[k].x * w.x + m[k].y * w.y) * inv, (m[k].y * w.x - m[k].x * w.y) * inv);
    double err = 0.0;
    for (int k = 0; k < 6; ++k) {
        const double2 u = mk<double2>(w.x * v[k].x - w.y * v[k].y, w.x * v[k].y + w.y * v[k].x);
        err = fmax(err, fmax(fabs(u.x - m[k].x), fabs(u.y - m[k].y)));
    }
    for (int c = 0; c < 3; ++c) {
        const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
        const double2 a1 = v[c1], b2 = v[3 + c2], a2 = v[c2], b1 = v[3 + c1];
        const double re = (a1.x * b2.x - a1.y * b2.y) - (a2.x * b1.x - a2.y * b1.y);
        const double im = (a1.x * b2.y + a1.y * b2.x) - (a2.x * b1.y + a2.y * b1.x);
        const double2 u = mk<double2>(w.x * re + w.y * im, w.y * re - w.x * im);
        err = fmax(err, fmax(fabs(u.x - m[6 + c].x), fabs(u.y - m[6 + c].y)));
    }
    return r > 0.0 && err <= tol * amax;
}

template <typename C2>
__global__ void k_canon_to_linkF12C(const C2 *canon, char *dst, uint32_t Vh, uint32_t nb64, int *fail)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)8 * Vh)
        return;
    const uint32_t qm = (uint32_t)(t / Vh), s = (uint32_t)(t - (size_t)qm * Vh);
    double2 v[6], w;
    if (!su3_factor_c(canon, qm, s, Vh, v, w, 1e-6))
        atomicAdd(fail, 1);
    char *b = dst + link_block_offset(qm / 4, qm % 4, s, nb64, LinkF12C::kBlock);
    const uint32_t l = s & 63;
    for (int c = 0; c < 3; ++c)
        reinterpret_cast<float4 *>(b + 1024 * c)[l] =
            make_float4((float)v[2 * c].x, (float)v[2 * c].y, (float)v[2 * c + 1].x, (float)v[2 * c + 1].y);
    reinterpret_cast<float2 *>(b + 3072)[l] = make_float2((float)w.x, (float)w.y);
}

__global__ void k_canon_to_linkH12C(const double2 *canon, char *dst, uint32_t Vh, uint32_t nb64, int *fail)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)8 * Vh)
        return;
    const uint32_t qm = (uint32_t)(t / Vh), s = (uint32_t)(t - (size_t)qm * Vh);
    double2 v[6], w;
    if (!su3_factor_c(canon, qm, s, Vh, v, w, 1e-6))
        atomicAdd(fail, 1);
    unsigned int h[6];
    for (int k = 0; k < 6; ++k)
        h[k] = pack_q2(v[k].x, v[k].y);
    char *b = dst + link_block_offset(qm / 4, qm % 4, s, nb64, LinkH12C::kBlock);
    const uint32_t l = s & 63;
    reinterpret_cast<uint4 *>(b)[l] = make_uint4(h[0], h[1], h[2], h[3]);
    reinterpret_cast<uint4 *>(b + 1024)[l] =
        make_uint4(h[4], h[5], __float_as_uint((float)w.x), __float_as_uint((float)w.y));
}

template <typename C2> static void canon_to_linkF12C_t(const C2 *canon, char *dst, uint32_t Vh, int *fail,
                                                       hipStream_t st)
{
    const size_t n = (size_t)8 * Vh;
    if (n == 0)
        return;
    hipLaunchKernelGGL(k_canon_to_linkF12C<C2>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, canon, dst, Vh,
                       nb64_of(Vh), fail);
    KQ_HIP_CHECK(hipGetLastError());
}

void canon_to_linkF12C(const double2 *canon, char *dst, uint32_t Vh, int *fail, hipStream_t st)
{
    canon_to_linkF12C_t(canon, dst, Vh, fail, st);
}

void canon_to_linkF12C(const float2 *canon, char *dst, uint32_t Vh, int *fail, hipStream_t st)
{
    canon_to_linkF12C_t(canon, dst, Vh, fail, st);
}

void canon_to_linkH12C(const double2 *canon, char *dst, uint32_t Vh, int *fail, hipStream_t st)
{
    const size_t n = (size_t)8 * Vh;
    if (n == 0)
        return;
    hipLaunchKernelGGL(k_canon_to_linkH12C, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, canon, dst, Vh,
                       nb64_of(Vh), fail);
    KQ_HIP_CHECK(hipGetLastError());
}

// LinkD14: w = principal cube root of det U (so V = U / w has det 1), rows 0 and 1 of V and w.
// A link counts in *fail when the reconstruction w (V0, V1, conj(V0 x V1)), formed as
// LinkD14::load forms it, differs from U by more than 1e-13 max|U_ij| in any entry: U is then not
// a complex multiple of an SU(3) matrix to FP64 accuracy (or is zero).
__global__ void k_canon_to_linkD14(const double2 *canon, char *dst, uint32_t Vh, uint32_t nb64, int *fail)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)8 * Vh)
        return;
    const uint32_t qm = (uint32_t)(t / Vh), s = (uint32_t)(t - (size_t)qm * Vh);
    double2 m[9];
    double amax = 0.0;
    for (int k = 0; k < 9; ++k) {
        m[k] = canon[((size_t)qm * 9 + k) * Vh + s];
        amax = fmax(amax, fmax(fabs(m[k].x), fabs(m[k].y)));
    }
    const double2 d = det3(m);
    const double r = cbrt(hypot(d.x, d.y)), ph = atan2(d.y, d.x) / 3.0;
    const double2 w = mk<double2>(r * cos(ph), r * sin(ph));
    const double inv = r > 0.0 ? 1.0 / (r * r) : 0.0;
    double2 v[6];
    for (int k = 0; k < 6; ++k)  // m / w = m conj(w) / |w|^2
        v[k] = mk<double2>((m[k].x * w.x + m[k].y * w.y) * inv, (m[k].y * w.x - m[k].x * w.y) * inv);
    double err = 0.0;
    for (int k = 0; k < 6; ++k) {
        const double2 u = mk<double2>(w.x * v[k].x - w.y * v[k].y, w.x * v[k].y + w.y * v[k].x);
        err = fmax(err, fmax(fabs(u.x - m[k].x), fabs(u.y - m[k].y)));
    }
    for (int c = 0; c < 3; ++c) {
        const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
        const double2 a1 = v[c1], b2 = v[3 + c2], a2 = v[c2], b1 = v[3 + c1];
        const double re = (a1.x * b2.x - a1.y * b2.y) - (a2.x * b1.x - a2.y * b1.y);
        const double im = (a1.x * b2.y + a1.y * b2.x) - (a2.x * b1.y + a2.y * b1.x);
        const double2 u = mk<double2>(w.x * re + w.y * im, w.y * re - w.x * im);
        err = fmax(err, fmax(fabs(u.x - m[6 + c].x), fabs(u.y - m[6 + c].y)));
    }
    if (!(r > 0.0 && err <= 1e-13 * amax))
        atomicAdd(fail, 1);
    double2 *b = reinterpret_cast<double2 *>(dst + link_block_offset(qm / 4, qm % 4, s, nb64, LinkD14::kBlock)) +
                 (s & 63);
    for (int k = 0; k < 6; ++k)
        b[64 * k] = v[k];
    b[64 * 6] = w;
}

void canon_to_linkD14(const double2 *canon, char *dst, uint32_t Vh, int *fail, hipStream_t st)
{
    const size_t n = (size_t)8 * Vh;
    if (n == 0)
        return;
    hipLaunchKernelGGL(k_canon_to_linkD14, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, canon, dst, Vh,
                       nb64_of(Vh), fail);
    KQ_HIP_CHECK(hipGetLastError());
}

void canon_to_linkH12(const double2 *canon, char *dst, uint32_t Vh, int *fail, hipStream_t st)
{
    const size_t n = (size_t)8 * Vh;
    if (n == 0)
        return;
    hipLaunchKernelGGL(k_canon_to_linkH12, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, canon, dst, Vh,
                       nb64_of(Vh), fail);
    KQ_HIP_CHECK(hipGetLastError());
}

template <typename C2>
__global__ void k_canon_to_linkF12(const C2 *canon, char *dst, uint32_t Vh, uint32_t nb64, int *fail)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)8 * Vh)
        return;
    const uint32_t qm = (uint32_t)(t / Vh), s = (uint32_t)(t - (size_t)qm * Vh);
    double2 v[9];
    double bsc;
    if (!su3_factor(canon, qm, s, Vh, v, bsc))
        atomicAdd(fail, 1);
    char *b = dst + link_block_offset(qm / 4, qm % 4, s, nb64, LinkF12::kBlock);
    const uint32_t l = s & 63;
    for (int c = 0; c < 3; ++c)
        reinterpret_cast<float4 *>(b + 1024 * c)[l] =
            make_float4((float)v[2 * c].x, (float)v[2 * c].y, (float)v[2 * c + 1].x, (float)v[2 * c + 1].y);
    reinterpret_cast<float *>(b + 3072)[l] = (float)bsc;
}

template <typename C2> static void canon_to_linkF12_t(const C2 *canon, char *dst, uint32_t Vh, int *fail, hipStream_t st)
{
    const size_t n = (size_t)8 * Vh;
    if (n == 0)
        return;
    hipLaunchKernelGGL(k_canon_to_linkF12<C2>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, canon, dst, Vh,
                       nb64_of(Vh), fail);
    KQ_HIP_CHECK(hipGetLastError());
}

void canon_to_linkF12(const double2 *canon, char *dst, uint32_t Vh, int *fail, hipStream_t st)
{
    canon_to_linkF12_t(canon, dst, Vh, fail, st);
}

void canon_to_linkF12(const float2 *canon, char *dst, uint32_t Vh, int *fail, hipStream_t st)
{
    canon_to_linkF12_t(canon, dst, Vh, fail, st);
}

template <typename HR, typename VR>
__global__ void k_vec_host_to_soa(const HR *raw, Vec<VR> v, int r)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= v.Vh)
        return;
    c2_t<VR> o[3];
    for (int c = 0; c < 3; ++c) {
        o[c].x = (VR)raw[6 * (size_t)s + 2 * c];
        o[c].y = (VR)raw[6 * (size_t)s + 2 * c + 1];
    }
    v.store(r, s, o);
}

template <typename HR, typename VR>
void vec_host_to_soa(const HR *raw, Vec<VR> v, int r, hipStream_t st)
{
    if (v.Vh == 0)
        return;
    hipLaunchKernelGGL((k_vec_host_to_soa<HR, VR>), dim3((v.Vh + 255) / 256), dim3(256), 0, st, raw, v, r);
    KQ_HIP_CHECK(hipGetLastError());
}

template <typename VR, typename HR>
__global__ void k_vec_soa_to_host(Vec<VR> v, int r, HR *raw)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= v.Vh)
        return;
    c2_t<VR> o[3];
    v.load(r, s, o);
    for (int c = 0; c < 3; ++c) {
        raw[6 * (size_t)s + 2 * c] = (HR)o[c].x;
        raw[6 * (size_t)s + 2 * c + 1] = (HR)o[c].y;
    }
}

template <typename VR, typename HR>
void vec_soa_to_host(Vec<VR> v, int r, HR *raw, hipStream_t st)
{
    if (v.Vh == 0)
        return;
    hipLaunchKernelGGL((k_vec_soa_to_host<VR, HR>), dim3((v.Vh + 255) / 256), dim3(256), 0, st, v, r, raw);
    KQ_HIP_CHECK(hipGetLastError());
}

template <typename XR, typename YR>
__global__ void k_convert_copy(const c2_t<XR> *x, c2_t<YR> *y, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        c2_t<YR> o;
        o.x = (YR)x[i].x;
        o.y = (YR)x[i].y;
        y[i] = o;
    }
}

template <typename XR, typename YR>
void convert_copy(const c2_t<XR> *x, c2_t<YR> *y, size_t n, hipStream_t st)
{
    if (n == 0)
        return;
    hipLaunchKernelGGL((k_convert_copy<XR, YR>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, y, n);
    KQ_HIP_CHECK(hipGetLastError());
}

// four 16-byte elements per thread, each wavefront-coalesced
__global__ void __launch_bounds__(256) k_copy16(const uint4 *__restrict__ a, uint4 *__restrict__ b, size_t n)
{
    const size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x;
    uint4 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (base + 256 * k < n)
            v[k] = a[base + 256 * k];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (base + 256 * k < n)
            b[base + 256 * k] = v[k];
}

void device_copy(const void *src, void *dst, size_t bytes, hipStream_t st)
{
    const size_t n = bytes / 16;
    if (n == 0)
        return;
    const unsigned nb = (unsigned)((n + 1023) / 1024);
    hipLaunchKernelGGL(k_copy16, dim3(nb), dim3(256), 0, st, static_cast<const uint4 *>(src), static_cast<uint4 *>(dst),
                       n);
    KQ_HIP_CHECK(hipGetLastError());
}

template void gauge_host_to_soa<double, double>(const double *, size_t, size_t, double2 *, uint32_t, double,
                                                hipStream_t);
template void gauge_host_to_soa<float, float>(const float *, size_t, size_t, float2 *, uint32_t, double, hipStream_t);
template void gauge_host_to_soa<float, double>(const float *, size_t, size_t, double2 *, uint32_t, double,
                                               hipStream_t);
template void canon_to_linkF<double>(const double2 *, char *, uint32_t, hipStream_t);
template void canon_to_linkF<float>(const float2 *, char *, uint32_t, hipStream_t);
template void vec_host_to_soa<double, double>(const double *, Vec<double>, int, hipStream_t);
template void vec_host_to_soa<float, float>(const float *, Vec<float>, int, hipStream_t);
template void vec_host_to_soa<double, float>(const double *, Vec<float>, int, hipStream_t);
template void vec_soa_to_host<double, double>(Vec<double>, int, double *, hipStream_t);
template void vec_soa_to_host<float, float>(Vec<float>, int, float *, hipStream_t);
template void vec_soa_to_host<float, double>(Vec<float>, int, double *, hipStream_t);
template void convert_copy<double, float>(const double2 *, float2 *, size_t, hipStream_t);
template void convert_copy<float, double>(const float2 *, double2 *, size_t, hipStream_t);
template void convert_copy<double, double>(const double2 *, double2 *, size_t, hipStream_t);
template void convert_copy<float, float>(const float2 *, float2 *, size_t, hipStream_t);

}  // namespace kq
