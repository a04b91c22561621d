// Device-resident multishift CG state and the kernels of one CGM iteration.
//
// One CgmState per right-hand side holds the Jegerlehner recurrence scalars of
// CGM::Invert (inverter_cgm.cpp:20-78). Every kernel of an iteration reads the
// state from device memory, so an iteration needs no host round trip and can
// be captured once in a hipGraph and replayed; a right-hand side that has
// converged (or hit max_iter) sets `stop`, after which the kernels skip it.
#pragma once

#include "kq_fields.h"

namespace kq {

constexpr int kMaxShift = 32;

template <typename S> struct CgmState {
    S rr;   // |r|^2 of the current residual
    S Bm1;  // beta_0 of the previous iteration
    S a[kMaxShift], B[kMaxShift], Z[kMaxShift], Zm1[kMaxShift];
    S shift[kMaxShift];  // reference convention: shift[0] = sigma_0, shift[k>=1] relative to it
    double tol;          // absolute squared-residual target (inverter_cgm.cpp:72-73)
    double r2max;        // |r|^2 at the last reliable update (mixed precision only)
    int nshift, maxTerm, maxTermUpd, iter, maxIter, stop, run, maxTermFold;
};

struct CgmLaunch {
    uint32_t Vh;
    uint32_t nblk;      // blocks of the BLAS site loops (= stride of part_b)
    uint32_t nblk_dsl;  // blocks of the hopping-term launch (= stride of part_a)
    int nrhs;        // right-hand sides in the batch
    int *rlist;      // compact list of live right-hand sides (written by k_cgm_begin)
    int *nact;       // its length
    int *any;        // 1 while any right-hand side is live
    double *part_a;  // [nrhs][nblk_dsl] partial sums of <p, A p>
    double *part_b;  // [nrhs][nblk] partial sums of |r|^2
};

template <typename S> void cgm_begin(CgmState<S> *st, const CgmLaunch &L, hipStream_t stream);
template <typename S> void cgm_scalars1(CgmState<S> *st, const CgmLaunch &L, hipStream_t stream);
template <typename S> void cgm_scalars2(CgmState<S> *st, const CgmLaunch &L, hipStream_t stream);
// r += beta_0 s ; partial |r|^2                        (inverter_cgm.cpp:50, 58)
template <typename S, typename VR>
void cgm_axpy_norm(CgmState<S> *st, Vec<VR> r, Vec<VR> s, const CgmLaunch &L, hipStream_t stream);
// x_k -= beta_k p_k ; p_k = zeta_k r + alpha_k p_k      (inverter_cgm.cpp:61-69)
// x and p: right-hand side r, shift k at base + (r*nshift + k)*3*Vh.
template <typename S, typename VR, typename XR>
void cgm_xp_update(CgmState<S> *st, XR *x, VR *p, Vec<VR> r, int nshift, const CgmLaunch &L, hipStream_t stream);
// state init: rr = |b|^2 from part_a, scalars as inverter_cgm.cpp:30-39
template <typename S>
void cgm_init_state(CgmState<S> *st, const CgmLaunch &L, hipStream_t stream);

// ---- reliable updates (mixed precision): systems listed in rel_list[0..*rel_n)
// r_hi = b - Ax (Ax = A_0 x_hi computed by the FP64 hopping term), r_lo = (VR) r_hi,
// part[r][blk] = |r_hi|^2; rr = r2max = |r_hi|^2 is then set by rel_set_rr.
template <typename VR>
void rel_residual(const double *b, const double *ax, Vec<VR> r_lo, const int *rel_list, const int *rel_n, int nsys,
                  double *part, uint32_t nblk, hipStream_t stream);
template <typename S>
void rel_set_rr(CgmState<S> *st, const int *rel_list, const int *rel_n, int nsys, const double *part, uint32_t nblk,
                hipStream_t stream);
// x_hi[r][k] += x_lo[r][k]; x_lo[r][k] = 0 for k < st[r].maxTermFold
template <typename S, typename VR>
void rel_fold(CgmState<S> *st, double *x_hi, VR *x_lo, int nshift, uint32_t Vh, const int *rel_list,
              const int *rel_n, int nsys, hipStream_t stream);

// part[k][blk] = |b - ax_k|^2 for k = list[0..nk) (or k < nk when list is null);
// b shared by all k, FP64
void resid2_partial(const double *b, const double *ax, int nk, const int *list, uint32_t Vh, double *part,
                    uint32_t nblk, hipStream_t stream);

// ---- generic BLAS-1 on colour vectors with deterministic partial sums
// part[r][blk] = sum over the block's sites of Re<x_r, y_r>
template <typename VR>
void vec_dot_partial(Vec<VR> x, Vec<VR> y, int nrhs, double *part, uint32_t nblk, hipStream_t stream);
// out[r] = sum_blk part[r][blk] (fixed order)
void reduce_partials(const double *part, uint32_t nblk, int nrhs, double *out, hipStream_t stream);

constexpr int kBlasNT = 256;
inline uint32_t blas_nblk(uint32_t Vh) { return (Vh + kBlasNT - 1) / kBlasNT; }

}  // namespace kq
