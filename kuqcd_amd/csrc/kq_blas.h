// BLAS-1 on flat arrays of complex numbers (whole colour-vector fields of the
// inner C++ API).
#pragma once

#include "kq_fields.h"

namespace kq {

// y[i] = a x[i] + b y[i] for i < n (b == 0: y = a x without reading y)
template <typename R> void flat_axpby(const c2_t<R> *x, c2_t<R> *y, size_t n, R a, R b, hipStream_t st);

// *out = sum_i Re(conj(x_i) y_i) = sum re*re + im*im, in FP64: one partial per workgroup of a
// fixed chunk of the array, added in a fixed order (part holds flat_dot_nblk(n) doubles)
constexpr uint32_t kFlatDotBlocks = 1024;
inline size_t flat_dot_chunk(size_t n)
{
    const size_t c = (n + kFlatDotBlocks - 1) / kFlatDotBlocks;
    return ((c + 255) / 256) * 256;
}
inline uint32_t flat_dot_nblk(size_t n)
{
    return n == 0 ? 1u : (uint32_t)((n + flat_dot_chunk(n) - 1) / flat_dot_chunk(n));
}
template <typename R>
void flat_dot(const c2_t<R> *x, const c2_t<R> *y, size_t n, double *part, double *out, hipStream_t st);

}  // namespace kq
