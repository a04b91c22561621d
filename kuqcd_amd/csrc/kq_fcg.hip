// Kernels of the flexible CG: fused vector updates with block partial sums,
// and single-block scalar kernels that finish the reductions in a fixed order
// and update the device-resident FcgState.
#include "kq_fcg.h"

namespace kq {

namespace {

constexpr int kNT = 256;

__device__ __forceinline__ bool halted(const FcgState *st) { return st->stop != 0; }

// block sum of (a, b) into part[2*blk], part[2*blk+1] (b ignored when part2 is false)
template <bool TWO>
__device__ __forceinline__ void block_partial(double a, double b, double *part)
{
    __shared__ double la[kNT / kWave], lb[kNT / kWave];
    a = wave_sum(a);
    if (TWO)
        b = wave_sum(b);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        la[threadIdx.x / kWave] = a;
        if (TWO)
            lb[threadIdx.x / kWave] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0.0, sb = 0.0;
        for (int w = 0; w < kNT / kWave; ++w) {
            sa += la[w];
            if (TWO)
                sb += lb[w];
        }
        if (TWO) {
            part[2 * blockIdx.x] = sa;
            part[2 * blockIdx.x + 1] = sb;
        } else {
            part[blockIdx.x] = sa;
        }
    }
}

// sum of part[k*stride + off], k < nblk, by one block in a fixed order; valid in thread 0
__device__ double block_total(const double *part, uint32_t nblk, int stride, int off)
{
    __shared__ double l[kNT];
    double s = 0.0;
    for (uint32_t k = threadIdx.x; k < nblk; k += kNT)
        s += part[(size_t)k * stride + off];
    l[threadIdx.x] = s;
    __syncthreads();
    for (int w = kNT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            l[threadIdx.x] += l[threadIdx.x + w];
        __syncthreads();
    }
    const double r = l[0];
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(kNT) k_fcg_dot(FcgState *st, const float2 *q, const float2 *z, size_t n,
                                                 double *part)
{
    if (halted(st))
        return;
    const size_t i = (size_t)blockIdx.x * kNT + threadIdx.x;
    double re = 0.0, im = 0.0;
    if (i < n) {
        const float2 a = q[i], b = z[i];
        re = (double)a.x * b.x + (double)a.y * b.y;
        im = (double)a.x * b.y - (double)a.y * b.x;
    }
    block_partial<true>(re, im, part);
}

__global__ void __launch_bounds__(kNT) k_fcg_beta(FcgState *st, const double *part, uint32_t nblk)
{
    if (halted(st))
        return;
    const double re = block_total(part, nblk, 2, 0);
    const double im = block_total(part, nblk, 2, 1);
    if (threadIdx.x == 0) {
        st->qz[0] = re;
        st->qz[1] = im;
        if (st->first || st->pq_prev == 0.0) {
            st->beta[0] = 0.0;
            st->beta[1] = 0.0;
        } else {
            st->beta[0] = re / st->pq_prev;
            st->beta[1] = im / st->pq_prev;
        }
    }
}

__global__ void __launch_bounds__(kNT) k_fcg_p_update(FcgState *st, const float2 *z, float2 *p, const float2 *r,
                                                      size_t n, double *part)
{
    if (halted(st))
        return;
    const size_t i = (size_t)blockIdx.x * kNT + threadIdx.x;
    double re = 0.0, im = 0.0;
    if (i < n) {
        const float br = (float)st->beta[0], bi = (float)st->beta[1];
        const float2 zi = z[i];
        float2 pi = mk<float2>(zi.x, zi.y);
        if (!st->first) {
            const float2 po = p[i];
            pi.x -= br * po.x - bi * po.y;
            pi.y -= br * po.y + bi * po.x;
        }
        p[i] = pi;
        const float2 ri = r[i];
        re = (double)pi.x * ri.x + (double)pi.y * ri.y;
        im = (double)pi.x * ri.y - (double)pi.y * ri.x;
    }
    block_partial<true>(re, im, part);
}

__global__ void __launch_bounds__(kNT) k_fcg_alpha(FcgState *st, const double *part_pq, uint32_t nblk_pq,
                                                   const double *part_pr, uint32_t nblk)
{
    if (halted(st))
        return;
    const double pq = block_total(part_pq, nblk_pq, 1, 0);
    const double prr = block_total(part_pr, nblk, 2, 0);
    const double pri = block_total(part_pr, nblk, 2, 1);
    if (threadIdx.x == 0) {
        st->pq = pq;
        st->pr[0] = prr;
        st->pr[1] = pri;
        const bool ok = pq > 0.0;
        st->alpha[0] = ok ? prr / pq : 0.0;
        st->alpha[1] = ok ? pri / pq : 0.0;
        st->pq_prev = pq;
        st->first = 0;
    }
}

__global__ void __launch_bounds__(kNT) k_fcg_xr_update(FcgState *st, float2 *x, float2 *r, const float2 *p,
                                                       const float2 *q, size_t n, double *part)
{
    if (halted(st))
        return;
    const size_t i = (size_t)blockIdx.x * kNT + threadIdx.x;
    double rr = 0.0;
    if (i < n) {
        const float ar = (float)st->alpha[0], ai = (float)st->alpha[1];
        const float2 pi = p[i], qi = q[i];
        float2 xi = x[i], ri = r[i];
        xi.x += ar * pi.x - ai * pi.y;
        xi.y += ar * pi.y + ai * pi.x;
        ri.x -= ar * qi.x - ai * qi.y;
        ri.y -= ar * qi.y + ai * qi.x;
        x[i] = xi;
        r[i] = ri;
        rr = (double)ri.x * ri.x + (double)ri.y * ri.y;
    }
    block_partial<false>(rr, 0.0, part);
}

__global__ void __launch_bounds__(kNT) k_fcg_check(FcgState *st, const double *part, uint32_t nblk)
{
    if (halted(st))
        return;
    const double rr = block_total(part, nblk, 1, 0);
    if (threadIdx.x == 0) {
        st->rr = rr;
        st->iter += 1;
        int reason = 0;
        if (!(rr > st->rr_stop))  // also stops on NaN
            reason = kFcgStopConverged;
        else if (rr <= st->rr_pause)
            reason = kFcgStopPause;
        else if (st->iter >= st->max_iter)
            reason = kFcgStopMaxIter;
        if (reason) {
            st->reason = reason;
            st->stop = 1;
            st->run = 0;
        }
    }
}

__global__ void k_fcg_fold(double2 *x64, float2 *x32, size_t n)
{
    const size_t i = (size_t)blockIdx.x * kNT + threadIdx.x;
    if (i < n) {
        const float2 a = x32[i];
        double2 b = x64[i];
        b.x += a.x;
        b.y += a.y;
        x64[i] = b;
        x32[i] = mk<float2>(0.f, 0.f);
    }
}

__global__ void __launch_bounds__(kNT) k_fcg_residual(const double2 *b, const double2 *ax, double2 *r64,
                                                      float2 *r32, size_t n, double *part)
{
    const size_t i = (size_t)blockIdx.x * kNT + threadIdx.x;
    double rr = 0.0;
    if (i < n) {
        const double2 bi = b[i], ai = ax[i];
        const double2 ri = mk<double2>(bi.x - ai.x, bi.y - ai.y);
        if (r64)
            r64[i] = ri;
        r32[i] = mk<float2>((float)ri.x, (float)ri.y);
        rr = ri.x * ri.x + ri.y * ri.y;
    }
    block_partial<false>(rr, 0.0, part);
}

__global__ void k_fcg_to_f32(const double2 *a, const double2 *b, double scale, float2 *out, size_t n)
{
    const size_t i = (size_t)blockIdx.x * kNT + threadIdx.x;
    if (i < n) {
        double2 v = a[i];
        if (b) {
            const double2 w = b[i];
            v.x -= w.x;
            v.y -= w.y;
        }
        out[i] = mk<float2>((float)(scale * v.x), (float)(scale * v.y));
    }
}

}  // namespace

uint32_t fcg_nblk(size_t n) { return (uint32_t)((n + kNT - 1) / kNT); }

#define KQ_FCG_LAUNCH(kernel, grid, ...)                                           \
    do {                                                                           \
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(kNT), 0, s, __VA_ARGS__);      \
        KQ_HIP_CHECK(hipGetLastError());                                           \
    } while (0)

void fcg_dot(FcgState *st, const float2 *q, const float2 *z, size_t n, double *part, hipStream_t s)
{
    KQ_FCG_LAUNCH(k_fcg_dot, fcg_nblk(n), st, q, z, n, part);
}
void fcg_beta(FcgState *st, const double *part, uint32_t nblk, hipStream_t s)
{
    KQ_FCG_LAUNCH(k_fcg_beta, 1, st, part, nblk);
}
void fcg_p_update(FcgState *st, const float2 *z, float2 *p, const float2 *r, size_t n, double *part, hipStream_t s)
{
    KQ_FCG_LAUNCH(k_fcg_p_update, fcg_nblk(n), st, z, p, r, n, part);
}
void fcg_alpha(FcgState *st, const double *part_pq, uint32_t nblk_pq, const double *part_pr, uint32_t nblk,
               hipStream_t s)
{
    KQ_FCG_LAUNCH(k_fcg_alpha, 1, st, part_pq, nblk_pq, part_pr, nblk);
}
void fcg_xr_update(FcgState *st, float2 *x, float2 *r, const float2 *p, const float2 *q, size_t n, double *part,
                   hipStream_t s)
{
    KQ_FCG_LAUNCH(k_fcg_xr_update, fcg_nblk(n), st, x, r, p, q, n, part);
}
void fcg_check(FcgState *st, const double *part, uint32_t nblk, hipStream_t s)
{
    KQ_FCG_LAUNCH(k_fcg_check, 1, st, part, nblk);
}
void fcg_fold(double2 *x64, float2 *x32, size_t n, hipStream_t s)
{
    if (n)
        KQ_FCG_LAUNCH(k_fcg_fold, fcg_nblk(n), x64, x32, n);
}
void fcg_residual(const double2 *b, const double2 *ax, double2 *r64, float2 *r32, size_t n, double *part,
                  hipStream_t s)
{
    if (n)
        KQ_FCG_LAUNCH(k_fcg_residual, fcg_nblk(n), b, ax, r64, r32, n, part);
}

void fcg_to_f32(const double2 *a, const double2 *b, double scale, float2 *out, size_t n, hipStream_t s)
{
    if (n)
        KQ_FCG_LAUNCH(k_fcg_to_f32, fcg_nblk(n), a, b, scale, out, n);
}

}  // namespace kq
