// Flexible CG, FCG(1), for one shifted system A = sigma - D D on one parity,
// with device-resident FP64 scalars so that an iteration can be captured in a
// hipGraph and replayed in chunks:
//
//   z = M r                      (preconditioner; z = r without one)
//   beta = <q, z> / <p, q>       (q, p of the previous iteration; 0 on the first)
//   p = z - beta p,  q = A p
//   alpha = <p, r> / <p, q>
//   x += alpha p,  r -= alpha q
//
// The A-orthogonalisation against the previous direction keeps the method a
// conjugate-direction method for a preconditioner that is neither Hermitian
// nor fixed. Vectors are FP32 (SoA colour-major, 3*Vh complex); every sum is
// an FP64 two-stage reduction in fixed order; the solution is accumulated in
// FP64 by reliable updates driven from the host.
#pragma once

#include "kq_dslash.h"

namespace kq {

struct FcgState {
    double qz[2];     // <q_prev, z>
    double pq;        // Re <p, q>
    double pq_prev;   // Re <p, q> of the previous iteration
    double pr[2];     // <p, r>
    double rr;        // |r|^2
    double alpha[2];  //
    double beta[2];   //
    double rr_pause;  // stop (for a reliable update) when rr <= rr_pause
    double rr_stop;   // stop when rr <= rr_stop
    double sigma;     // shift of A
    int iter;         // iterations done
    int max_iter;     // stop when iter reaches it
    int first;        // next iteration starts a new direction sequence (beta = 0)
    int stop;         // 1: every kernel of the iteration does nothing
    int run;          // !stop, for the hopping-term kernels
    int reason;       // why it stopped: 1 rr_pause, 2 rr_stop, 3 max_iter
};

constexpr int kFcgStopPause = 1, kFcgStopConverged = 2, kFcgStopMaxIter = 3;

uint32_t fcg_nblk(size_t n);  // blocks of the vector kernels for n complex entries

// partials of <q, z> (complex) -> part[2*blk], part[2*blk+1]
void fcg_dot(FcgState *st, const float2 *q, const float2 *z, size_t n, double *part, hipStream_t s);
// qz = sum(part); beta = first ? 0 : qz / pq_prev
void fcg_beta(FcgState *st, const double *part, uint32_t nblk, hipStream_t s);
// p = z - beta p (first: p = z); partials of <p, r>
void fcg_p_update(FcgState *st, const float2 *z, float2 *p, const float2 *r, size_t n, double *part, hipStream_t s);
// pq = sum(part_pq) (real, [nblk_pq]); pr = sum(part_pr) (complex); alpha = pr / pq; first = 0
void fcg_alpha(FcgState *st, const double *part_pq, uint32_t nblk_pq, const double *part_pr, uint32_t nblk,
               hipStream_t s);
// x += alpha p; r -= alpha q; partials of |r|^2
void fcg_xr_update(FcgState *st, float2 *x, float2 *r, const float2 *p, const float2 *q, size_t n, double *part,
                   hipStream_t s);
// rr = sum(part); iter += 1; stop on rr_stop, rr_pause or max_iter
void fcg_check(FcgState *st, const double *part, uint32_t nblk, hipStream_t s);

// reliable update, outside the graph: x64 += x32, x32 = 0
void fcg_fold(double2 *x64, float2 *x32, size_t n, hipStream_t s);
// r64 = b - ax (r64 may be null), r32 = r; partials of |r|^2 (real, [nblk])
void fcg_residual(const double2 *b, const double2 *ax, double2 *r64, float2 *r32, size_t n, double *part,
                  hipStream_t s);
// out = (float)(scale * (a - b)); b may be null
void fcg_to_f32(const double2 *a, const double2 *b, double scale, float2 *out, size_t n, hipStream_t s);

}  // namespace kq
