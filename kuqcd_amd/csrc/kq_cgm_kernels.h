// Device-resident multishift CG state and the kernels of one CGM iteration.
//
// One CgmState per system (a right-hand side with its list of shifts) holds the
// Jegerlehner recurrence scalars of CGM::Invert (inverter_cgm.cpp:20-78).
// Every kernel of an iteration reads the states from device memory, so an
// iteration needs no host round trip and is captured once in a hipGraph and
// replayed; a system that has converged (or hit max_iter) sets `stop`, and one
// that needs a reliable update may set `pause`; the kernels skip both.
//
// A system's search directions p_k and solutions x_k live at element offset
// off[sys] + k*3*Vh of their arrays (complex units), so systems with different
// numbers of shifts share one pool.
#pragma once

#include "kq_fields.h"

namespace kq {

constexpr int kMaxShift = 32;

template <typename S> struct CgmState {
    S rr;   // |r|^2 of the current residual
    S Bm1;  // beta_0 of the previous iteration
    S a[kMaxShift], B[kMaxShift], Z[kMaxShift], Zm1[kMaxShift];
    S shift[kMaxShift];       // reference convention: shift[0] = sigma_0, shift[k>=1] relative to it
    double tolk[kMaxShift];   // absolute squared-residual target per shift (inverter_cgm.cpp:72-73)
    double r2max;             // |r|^2 at the last reliable update (mixed precision)
    double pause2;            // pause when |r|^2 < pause2 * r2max (0: never)
    int nshift, maxTerm, maxTermUpd, iter, maxIter, stop, run, maxTermFold, pause;
    int discard;  // set by the host for one fold: drop x_lo instead of adding it (diverged iteration)
    // deferred solution updates (multishift systems of the mixed solver): history slot of the
    // system (-1: x and p of every shift are updated each iteration) and the number of
    // iterations recorded since the last segment fold
    int dslot, hist;
};

// Deferred solution updates. In CGM::Invert every live shift k updates
//   x_k -= B_k p_k ;  p_k = Z_k r + a_k p_k            (inverter_cgm.cpp:61-69)
// each iteration: 4 vector passes per shift. Over a segment of J iterations every p_k is
// p_k(start) plus a combination of the residuals r_1..r_J of the segment, so a system with a
// history slot keeps only p_0 (the hopping term's source) current and records r and the
// scalars B_k, Z_k, a_k of each iteration; a segment fold then adds each shift's J updates to
// its FP64 solution in one pass and rebuilds p_k (k >= 1) from p_k(start) and the recorded
// residuals. A fold happens whenever the solver folds anyway (reliable update, dropped shift,
// stop) and at the latest after kHistMax iterations.
constexpr int kHistMax = 32;
// recorded scalars per history slot and iteration: B_k, Z_k (after the iteration), a_k, and
// maxTermUpd in field 3
constexpr int kHistFields = 4;
__host__ __device__ inline size_t hist_scalar_index(int dslot, int j, int field, int k)
{
    return (((size_t)dslot * kHistMax + j) * kHistFields + field) * kMaxShift + k;
}

struct CgmLaunch {
    uint32_t Vh;
    uint32_t nblk;      // blocks of the BLAS site loops (= stride of part_b)
    uint32_t nblk_dsl;  // blocks of the hopping-term launch (= stride of part_a)
    int nsys;           // system slots (k_cgm_begin scans all of them)
    int nlaunch;        // slots of the live list the per-system kernels cover (>= live systems)
    int *rlist;         // compact list of live systems (written by k_cgm_begin)
    int *nact;          // its length
    int *any;           // 1 while any system is live
    double *part_a;     // [nsys][nblk_dsl] partial sums of <p, A p>
    double *part_b;     // [nsys][nblk] partial sums of |r|^2
    const size_t *off;  // [nsys] element offset of each system's p/x block
    double *shist;      // recorded scalars (hist_scalar_index), null without deferred updates
    void *rhist;        // recorded residuals: vector (dslot * kHistMax + j), 3 * Vh complex each
};

template <typename S> void cgm_begin(CgmState<S> *st, const CgmLaunch &L, hipStream_t stream);
template <typename S> void cgm_scalars1(CgmState<S> *st, const CgmLaunch &L, hipStream_t stream);
template <typename S> void cgm_scalars2(CgmState<S> *st, const CgmLaunch &L, hipStream_t stream);
// r += beta_0 s ; partial |r|^2                        (inverter_cgm.cpp:50, 58)
template <typename S, typename VR>
void cgm_axpy_norm(CgmState<S> *st, Vec<VR> r, Vec<VR> s, const CgmLaunch &L, hipStream_t stream);
// x_k -= beta_k p_k ; p_k = zeta_k r + alpha_k p_k      (inverter_cgm.cpp:61-69)
template <typename S, typename VR, typename XR>
void cgm_xp_update(CgmState<S> *st, XR *x, VR *p, Vec<VR> r, const CgmLaunch &L, hipStream_t stream);

// ---- deferred solution updates (segment fold) of the listed systems with a history slot:
// coefficients of each shift k < maxTermFold from the recorded scalars (coef: per slot and
// shift, [c_p, c_r[0..kHistMax)] for x and [d_p, d_r[0..kHistMax)] for p), then
//   x_hi[k] += c_p p_k(start) + sum_j c_r[j] r_j   (if fold_keeps(discard, k, maxTerm))
//   p_k      = d_p p_k(start) + sum_j d_r[j] r_j   (1 <= k < maxTerm)
// with p_0(start) in p0snap[dslot]; then p0snap = p_0 and hist = 0 (seg_reset, also at the
// start and restart of a system). kmax bounds maxTermFold over the listed systems.
constexpr int kCoefStride = 2 * (kHistMax + 1);
// One coefficient pair of shift k over the J recorded iterations of one history slot (h points
// at the slot's first scalar, hist_scalar_index(dslot, 0, 0, 0)): for m = -1 the start vector
// p_k(start), for 0 <= m < kHistMax the recorded residual r_m. cx is its coefficient in the
// segment's update of x_k, cp its coefficient in p_k at the end of the segment. Shift k was
// updated by the first Jk iterations (k < maxTermUpd_j); within them
//   x_k += -B_j p_k,j ,   p_k,j+1 = Z_j r_j + a_j p_k,j ,
// so a term that enters p_k with weight w (1 for p_k(start) at j = 0, Z_m for r_m at j = m + 1)
// adds -B_j w to cx at every later iteration j and is scaled by a_j. keep_x = false (a
// diverged segment) drops the x part.
__host__ __device__ inline void seg_coefficient(const double *h, int J, int k, int m, bool keep_x, double &cx,
                                                double &cp)
{
    const size_t step = (size_t)kHistFields * kMaxShift;
    int Jk = 0;
    while (Jk < J && Jk < kHistMax && k < (int)h[(size_t)Jk * step + 3 * kMaxShift])
        ++Jk;
    cx = 0.0;
    cp = 0.0;
    if (m >= Jk)
        return;  // r_m came after shift k was dropped
    // the start vector is in p_k,0 (iterations from 0 on see it); r_m enters p_k,m+1
    double w = (m < 0) ? 1.0 : h[(size_t)m * step + kMaxShift + k];
    double sx = 0.0;
    for (int j = m + 1; j < Jk; ++j) {
        sx -= h[(size_t)j * step + k] * w;
        w *= h[(size_t)j * step + 2 * kMaxShift + k];
    }
    cx = keep_x ? sx : 0.0;
    cp = w;
}

template <typename S>
void seg_coef(const CgmState<S> *st, const double *shist, double *coef, const int *list, const int *n, int nsys,
              hipStream_t stream);
// XR: real type of the solutions the folds add to (double: x_hi, or the FP64 solutions; float:
// the FP32 solutions of the float reference method)
template <typename S, typename VR, typename XR>
void seg_fold(const CgmState<S> *st, const double *coef, XR *x, VR *p, const VR *p0snap, const VR *rhist,
              const size_t *off, uint32_t Vh, const int *list, const int *n, int nsys, int kmax, hipStream_t stream);
template <typename S, typename VR>
void seg_reset(CgmState<S> *st, const VR *p, VR *p0snap, const size_t *off, uint32_t Vh, const int *list, const int *n,
               int nsys, hipStream_t stream);

// ---- start of listed systems (list[0..*n)): rr = r2max = sum of part_b,
// scalars as inverter_cgm.cpp:30-39, stop = pause = 0 (stop = 1 if nshift <= 0)
template <typename S>
void cgm_init_list(CgmState<S> *st, const int *list, const int *n, int nsys, const double *part, uint32_t nblk,
                   hipStream_t stream);

// ---- reliable updates (mixed precision), listed systems
// r_hi = b - Ax (Ax = A_0 x_hi by the FP64 hopping term), r_lo = (VR) r_hi,
// part[sys][blk] = |r_hi|^2; rel_set_rr then sets rr = r2max = |r_hi|^2.
template <typename VR>
void rel_residual(const double *b, const double *ax, Vec<VR> r_lo, const int *list, const int *n, int nsys,
                  double *part, uint32_t nblk, hipStream_t stream);
template <typename S>
void rel_set_rr(CgmState<S> *st, const int *list, const int *n, int nsys, const double *part, uint32_t nblk,
                hipStream_t stream);
// restart of listed systems that have a single live shift (after a reliable update
// has set r = b - A x): the recurrence starts afresh from r (p = r is copied by
// copy_r_to_p) while iter, maxTerm and the folded solutions are kept
template <typename S>
void cgm_restart_list(CgmState<S> *st, const int *list, const int *n, int nsys, hipStream_t stream);

// copy `bytes` (a multiple of 8) with a kernel; either side may be mapped pinned host
// memory, which avoids a copy-engine round trip for the small state transfers
void copy_words(const void *src, void *dst, size_t bytes, hipStream_t stream);

// list[0..n) = src[0..n), *count = n (src may be mapped pinned host memory)
void set_int_list(const int *src, int *list, int *count, int n, hipStream_t stream);
// Whether the fold adds x_lo[k] to x_hi[k]. A system whose iteration diverged
// (discard set by the host) drops the unfolded updates of the shifts that are still
// iterating (k < maxTerm); the shifts that converged and were dropped since the last
// fold (maxTerm <= k < maxTermFold) finished before the divergence was seen, and keep
// their updates.
__host__ __device__ inline bool fold_keeps(int discard, int k, int max_term)
{
    return discard == 0 || k >= max_term;
}
// x_hi[k] += x_lo[k] (if fold_keeps); x_lo[k] = 0 for k < st[sys].maxTermFold
template <typename S, typename VR>
void rel_fold(CgmState<S> *st, double *x_hi, VR *x_lo, const size_t *off, uint32_t Vh, const int *list, const int *n,
              int nsys, hipStream_t stream);
// p_k = r for k < st[sys].nshift (start of a system)
template <typename S, typename VR>
void copy_r_to_p(const CgmState<S> *st, Vec<VR> r, VR *p, const size_t *off, const int *list, const int *n, int nsys,
                 hipStream_t stream);

// part[k][blk] = |b - ax_k|^2 for k = list[0..nk) (or k < nk when list is null);
// b shared by all k, FP64
void resid2_partial(const double *b, const double *ax, int nk, const int *list, uint32_t Vh, double *part,
                    uint32_t nblk, hipStream_t stream);

// ---- generic BLAS-1 on colour vectors with deterministic partial sums
// part[r][blk] = sum over the block's sites of Re<x_r, y_r> for r = list[..] (all r < nrhs if null)
template <typename VR>
void vec_dot_partial(Vec<VR> x, Vec<VR> y, int nrhs, double *part, uint32_t nblk, hipStream_t stream,
                     const int *list = nullptr, const int *n = nullptr);
// out[r] = sum_blk part[r][blk] (fixed order)
void reduce_partials(const double *part, uint32_t nblk, int nrhs, double *out, hipStream_t stream);

constexpr int kBlasNT = 256;
inline uint32_t blas_nblk(uint32_t Vh) { return (Vh + kBlasNT - 1) / kBlasNT; }

}  // namespace kq
