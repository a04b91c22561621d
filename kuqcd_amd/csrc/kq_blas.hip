// BLAS-1 on whole colour-vector fields (flat arrays of complex numbers) for the
// inner C++ API: y = a x + b y, and Re<x, y> with one FP64 partial sum per
// workgroup added in a fixed order (spinor_field.cpp:130-252).
#include "kq_blas.h"

namespace kq {

// y = a x + b y (b == 0: y = a x, y not read); arithmetic in R as the reference's
// loops over T (spinor_field.cpp:130-190)
template <typename R>
__global__ void __launch_bounds__(256) k_axpby_flat(const c2_t<R> *__restrict__ x, c2_t<R> *__restrict__ y, size_t n,
                                                    R a, R b)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const c2_t<R> xv = x[i];
        if (b == R(0)) {
            y[i] = mk<c2_t<R>>(a * xv.x, a * xv.y);
        } else {
            const c2_t<R> yv = y[i];
            y[i] = mk<c2_t<R>>(a * xv.x + b * yv.x, a * xv.y + b * yv.y);
        }
    }
}

template <typename R> void flat_axpby(const c2_t<R> *x, c2_t<R> *y, size_t n, R a, R b, hipStream_t st)
{
    if (n == 0)
        return;
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_axpby_flat<R>, dim3(grid), dim3(256), 0, st, x, y, n, a, b);
    KQ_HIP_CHECK(hipGetLastError());
}

// part[blk] = sum over the block's chunk of re x re y + im x im y, in FP64
template <typename R>
__global__ void __launch_bounds__(256) k_dot_flat(const c2_t<R> *__restrict__ x, const c2_t<R> *__restrict__ y,
                                                  size_t n, size_t chunk, double *part)
{
    __shared__ double lds[256];
    const size_t i0 = (size_t)blockIdx.x * chunk, i1 = (i0 + chunk < n) ? i0 + chunk : n;
    double acc = 0.0;
    for (size_t i = i0 + threadIdx.x; i < i1; i += 256) {
        const c2_t<R> a = x[i], b = y[i];
        acc += (double)a.x * (double)b.x + (double)a.y * (double)b.y;
    }
    lds[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            lds[threadIdx.x] += lds[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        part[blockIdx.x] = lds[0];
}

__global__ void __launch_bounds__(256) k_sum_flat(const double *part, uint32_t n, double *out)
{
    __shared__ double lds[256];
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += 256)
        acc += part[i];
    lds[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            lds[threadIdx.x] += lds[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        *out = lds[0];
}

template <typename R>
void flat_dot(const c2_t<R> *x, const c2_t<R> *y, size_t n, double *part, double *out, hipStream_t st)
{
    const size_t chunk = flat_dot_chunk(n);
    const uint32_t nblk = flat_dot_nblk(n);
    if (n > 0) {
        hipLaunchKernelGGL(k_dot_flat<R>, dim3(nblk), dim3(256), 0, st, x, y, n, chunk, part);
        KQ_HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_sum_flat, dim3(1), dim3(256), 0, st, part, n > 0 ? nblk : 0u, out);
    KQ_HIP_CHECK(hipGetLastError());
}

template void flat_axpby<double>(const double2 *, double2 *, size_t, double, double, hipStream_t);
template void flat_axpby<float>(const float2 *, float2 *, size_t, float, float, hipStream_t);
template void flat_dot<double>(const double2 *, const double2 *, size_t, double *, double *, hipStream_t);
template void flat_dot<float>(const float2 *, const float2 *, size_t, double *, double *, hipStream_t);

}  // namespace kq
