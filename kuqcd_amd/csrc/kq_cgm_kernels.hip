// Kernels of one multishift CG iteration (CGM::Invert, inverter_cgm.cpp:42-78)
// with the recurrence scalars kept in device memory, plus the start-up and
// reliable-update kernels of the mixed-precision solver.
//
// Per iteration:  begin -> D (p0 -> tmp) -> D + axmy + <p0,s> -> scalars1 ->
//                 r += b0 s, |r|^2 -> scalars2 -> x/p update of every live shift
// Reductions are two-stage and deterministic: each workgroup writes one FP64
// partial sum, and one workgroup adds the partials in a fixed order.
#include "kq_cgm_kernels.h"

namespace kq {

// ------------------------------------------------------------ reduce helpers
// Sum part[0..n) in a fixed order with one workgroup of 256 threads.
__device__ inline double block_reduce_fixed(const double *part, uint32_t n)
{
    __shared__ double lds[256];
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
        acc += part[i];
    lds[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            lds[threadIdx.x] += lds[threadIdx.x + w];
        __syncthreads();
    }
    return lds[0];
}

// ------------------------------------------------------------ begin
template <typename S> __global__ void k_cgm_begin(CgmState<S> *st, int nsys, int *rlist, int *nact, int *any)
{
    if (threadIdx.x != 0)
        return;
    int n = 0;
    for (int r = 0; r < nsys; ++r) {
        const int live = (st[r].stop == 0) && (st[r].pause == 0);
        st[r].run = live;
        if (live)
            rlist[n++] = r;
    }
    *nact = n;
    *any = n > 0;
}

template <typename S> void cgm_begin(CgmState<S> *st, const CgmLaunch &L, hipStream_t stream)
{
    hipLaunchKernelGGL(k_cgm_begin<S>, dim3(1), dim3(64), 0, stream, st, L.nsys, L.rlist, L.nact, L.any);
    KQ_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------ scalars 1
// pAp = <p0, s>;  beta_0 = -|r|^2 / pAp;  zeta/beta recurrences of the shifted
// systems (inverter_cgm.cpp:47-57). The FP contraction is switched off so the
// recurrence rounds exactly as written; the double literal 1.0 of the
// reference keeps its type (the denominator is formed in double for S=float).
template <typename S>
__global__ void __launch_bounds__(256) k_cgm_scalars1(CgmState<S> *st, const int *rlist, const int *nact,
                                                      const double *part, uint32_t nblk, double *shist)
{
#pragma clang fp contract(off)
    const int slot = blockIdx.x;
    if (*nact <= slot)
        return;
    const int r = rlist[slot];
    const double pAp_d = block_reduce_fixed(part + (size_t)r * nblk, nblk);
    CgmState<S> &c = st[r];
    // every thread reads the values the shifted updates share before any thread writes
    const S B0 = -c.rr / (S)pAp_d;
    const S Bm1 = c.Bm1, a0 = c.a[0];
    const int mt = c.maxTerm;
    const bool rec = shist != nullptr && c.dslot >= 0 && c.hist < kHistMax;
    double *h = rec ? shist + hist_scalar_index(c.dslot, c.hist, 0, 0) : nullptr;
    __syncthreads();
    // one shifted system per thread; each touches only its own entries
    for (int j = 1 + (int)threadIdx.x; j < mt; j += blockDim.x) {
        const S zm1 = c.Zm1[j], z = c.Z[j];
        const S num = Bm1 * zm1;
        const S t1 = B0 * a0 * (zm1 - z);
        const S t2 = zm1 * Bm1;
        const S t3 = c.shift[j] * B0;
        const S rr1 = (S)((double)num / ((double)t1 + (double)t2 * (1.0 - (double)t3)));
        const S znew = z * rr1, bnew = B0 * rr1;
        c.Zm1[j] = z;
        c.Z[j] = znew;
        c.B[j] = bnew;
        if (rec) {
            h[j] = (double)bnew;
            h[kMaxShift + j] = (double)znew;
        }
    }
    if (threadIdx.x == 0) {
        c.B[0] = B0;
        c.Bm1 = B0;
        if (rec) {
            h[0] = (double)B0;
            h[kMaxShift] = (double)c.Z[0];
        }
    }
}

template <typename S> void cgm_scalars1(CgmState<S> *st, const CgmLaunch &L, hipStream_t stream)
{
    hipLaunchKernelGGL(k_cgm_scalars1<S>, dim3(L.nlaunch), dim3(256), 0, stream, st, L.rlist, L.nact, L.part_a,
                       L.nblk_dsl, L.shist);
    KQ_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------ scalars 2
// |r|^2 new, alpha recurrences, drop converged shifts from the end,
// iteration count and stop flag (inverter_cgm.cpp:58-78); pause flag for the
// reliable-update schedule.
template <typename S>
__global__ void __launch_bounds__(256) k_cgm_scalars2(CgmState<S> *st, const int *rlist, const int *nact,
                                                      const double *part, uint32_t nblk, double *shist)
{
#pragma clang fp contract(off)
    const int slot = blockIdx.x;
    if (*nact <= slot)
        return;
    const int r = rlist[slot];
    const double l2_d = block_reduce_fixed(part + (size_t)r * nblk, nblk);
    CgmState<S> &c = st[r];
    // every thread reads the values the shifted updates share before thread 0 writes them
    const S lambda2 = (S)l2_d;
    const S a0 = lambda2 / c.rr;
    const S B0 = c.B[0];
    const int mt = c.maxTerm;
    const bool rec = shist != nullptr && c.dslot >= 0 && c.hist < kHistMax;
    double *h = rec ? shist + hist_scalar_index(c.dslot, c.hist, 2, 0) : nullptr;
    __syncthreads();
    for (int j = 1 + (int)threadIdx.x; j < mt; j += blockDim.x) {
        const S aj = a0 * c.Z[j] * c.B[j] / (c.Zm1[j] * B0);
        c.a[j] = aj;
        if (rec)
            h[j] = (double)aj;
    }
    if (threadIdx.x != 0)
        return;
    c.a[0] = a0;
    c.rr = lambda2;
    c.maxTermUpd = mt;
    if (rec) {
        h[0] = (double)a0;
        h[kMaxShift] = (double)mt;
        c.hist += 1;
    }
    // drop converged shifts from the end of the list (inverter_cgm.cpp:71-76)
    while (c.maxTerm > 0) {
        const S lambda = c.Z[c.maxTerm - 1] * c.Z[c.maxTerm - 1] * lambda2;
        if (!((double)lambda < c.tolk[c.maxTerm - 1]))
            break;
        c.maxTerm--;
    }
    c.iter += 1;
    if (c.maxTerm <= 0 || c.iter >= c.maxIter)
        c.stop = 1;
    if (c.pause2 > 0.0 && l2_d < c.pause2 * c.r2max)
        c.pause = 1;
}

template <typename S> void cgm_scalars2(CgmState<S> *st, const CgmLaunch &L, hipStream_t stream)
{
    hipLaunchKernelGGL(k_cgm_scalars2<S>, dim3(L.nlaunch), dim3(256), 0, stream, st, L.rlist, L.nact, L.part_b,
                       L.nblk, L.shist);
    KQ_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------ r += b0 s, |r|^2
template <typename S, typename VR>
__global__ void __launch_bounds__(kBlasNT) k_axpy_norm(const CgmState<S> *st, Vec<VR> r, Vec<VR> s,
                                                       const int *rlist, const int *nact, double *part,
                                                       uint32_t nblk, c2_t<VR> *rhist)
{
    using V2 = c2_t<VR>;
    __shared__ double lds[kBlasNT / kWave];
    const int slot = blockIdx.y;
    if (*nact <= slot)
        return;
    const int rh = rlist[slot];
    const VR b0 = (VR)st[rh].B[0];
    const uint32_t i = blockIdx.x * kBlasNT + threadIdx.x;
    double acc = 0.0;
    if (i < r.Vh) {
        V2 rv[3], sv[3];
        r.load(rh, i, rv);
        s.load(rh, i, sv);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            rv[c].x += b0 * sv[c].x;
            rv[c].y += b0 * sv[c].y;
            acc += (double)rv[c].x * (double)rv[c].x + (double)rv[c].y * (double)rv[c].y;
        }
        r.store(rh, i, rv);
        const int ds = st[rh].dslot, hj = st[rh].hist;
        if (rhist != nullptr && ds >= 0 && hj < kHistMax) {
            V2 *h = rhist + ((size_t)ds * kHistMax + hj) * 3 * r.Vh + i;
#pragma unroll
            for (int c = 0; c < 3; ++c)
                st_nt(h + (size_t)c * r.Vh, rv[c]);
        }
    }
    const double tot = block_sum<kBlasNT>(acc, lds);
    if (threadIdx.x == 0)
        part[(size_t)rh * nblk + blockIdx.x] = tot;
}

template <typename S, typename VR>
void cgm_axpy_norm(CgmState<S> *st, Vec<VR> r, Vec<VR> s, const CgmLaunch &L, hipStream_t stream)
{
    hipLaunchKernelGGL((k_axpy_norm<S, VR>), dim3(L.nblk, L.nlaunch), dim3(kBlasNT), 0, stream, st, r, s, L.rlist,
                       L.nact, L.part_b, L.nblk, static_cast<c2_t<VR> *>(L.rhist));
    KQ_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------ x/p update
template <typename S, typename VR, typename XR>
__global__ void __launch_bounds__(kBlasNT) k_xp_update(const CgmState<S> *st, XR *x, VR *p, Vec<VR> r,
                                                       const size_t *off, const int *rlist, const int *nact)
{
    using V2 = c2_t<VR>;
    using X2 = c2_t<XR>;
    const int slot = blockIdx.y;
    if (*nact <= slot)
        return;
    const int rh = rlist[slot];
    const uint32_t i = blockIdx.x * kBlasNT + threadIdx.x;
    const uint32_t Vh = r.Vh;
    if (i >= Vh)
        return;
    const CgmState<S> &c = st[rh];
    // a system with deferred solution updates keeps only p_0 current (its x_k and p_k, k >= 1,
    // are brought up to date by the next segment fold)
    const bool defer = c.dslot >= 0;
    const int m = defer ? 1 : c.maxTermUpd;
    const size_t o = off[rh];
    V2 rv[3];
    r.load(rh, i, rv);
    for (int k = 0; k < m; ++k) {
        const size_t base = o + (size_t)k * 3 * Vh + i;
        const XR bk = (XR)c.B[k];
        const VR zk = (VR)c.Z[k], ak = (VR)c.a[k];
        V2 *pp = reinterpret_cast<V2 *>(p) + base;
        X2 *xx = reinterpret_cast<X2 *>(x) + base;
        // streaming accesses except the store of the unshifted direction p_0,
        // which the next hopping term gathers: the solutions and the shifted
        // directions are not touched again before the next update
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
            const V2 pv = ld_nt(pp + (size_t)cc * Vh);
            if (!defer) {
                X2 xv = ld_nt(xx + (size_t)cc * Vh);
                xv.x -= bk * (XR)pv.x;
                xv.y -= bk * (XR)pv.y;
                st_nt(xx + (size_t)cc * Vh, xv);
            }
            V2 np;
            np.x = zk * rv[cc].x + ak * pv.x;
            np.y = zk * rv[cc].y + ak * pv.y;
            if (k == 0)
                pp[(size_t)cc * Vh] = np;
            else
                st_nt(pp + (size_t)cc * Vh, np);
        }
    }
}

template <typename S, typename VR, typename XR>
void cgm_xp_update(CgmState<S> *st, XR *x, VR *p, Vec<VR> r, const CgmLaunch &L, hipStream_t stream)
{
    hipLaunchKernelGGL((k_xp_update<S, VR, XR>), dim3(L.nblk, L.nlaunch), dim3(kBlasNT), 0, stream, st, x, p, r, L.off,
                       L.rlist, L.nact);
    KQ_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------ start of systems
template <typename S>
__global__ void __launch_bounds__(256) k_cgm_init_list(CgmState<S> *st, const int *list, const int *n,
                                                       const double *part, uint32_t nblk)
{
    const int slot = blockIdx.x;
    if (*n <= slot)
        return;
    const int r = list[slot];
    const double rr = block_reduce_fixed(part + (size_t)r * nblk, nblk);
    if (threadIdx.x != 0)
        return;
    CgmState<S> &c = st[r];
    c.rr = (S)rr;
    c.Bm1 = (S)1.0;
    for (int k = 0; k < kMaxShift; ++k) {
        c.a[k] = (S)0.0;
        c.B[k] = (S)1.0;
        c.Z[k] = (S)1.0;
        c.Zm1[k] = (S)1.0;
    }
    c.maxTerm = c.nshift;
    c.maxTermUpd = c.nshift;
    c.maxTermFold = c.nshift;
    c.r2max = rr;
    c.iter = 0;
    c.pause = 0;
    c.discard = 0;
    c.hist = 0;
    // the reference's do-while runs at least one iteration even for max_iter <= 0
    c.stop = (c.nshift <= 0) ? 1 : 0;
    c.run = 0;
}

template <typename S>
void cgm_init_list(CgmState<S> *st, const int *list, const int *n, int nsys, const double *part, uint32_t nblk,
                   hipStream_t stream)
{
    hipLaunchKernelGGL(k_cgm_init_list<S>, dim3(nsys), dim3(256), 0, stream, st, list, n, part, nblk);
    KQ_HIP_CHECK(hipGetLastError());
}

template <typename S, typename VR>
__global__ void __launch_bounds__(kBlasNT) k_copy_r_to_p(const CgmState<S> *st, Vec<VR> r, VR *p, const size_t *off,
                                                         const int *list, const int *n)
{
    using V2 = c2_t<VR>;
    const int slot = blockIdx.y;
    if (*n <= slot)
        return;
    const int rh = list[slot];
    const uint32_t i = blockIdx.x * kBlasNT + threadIdx.x;
    const uint32_t Vh = r.Vh;
    if (i >= Vh)
        return;
    V2 rv[3];
    r.load(rh, i, rv);
    const int m = st[rh].nshift;
    for (int k = 0; k < m; ++k) {
        V2 *pp = reinterpret_cast<V2 *>(p) + off[rh] + (size_t)k * 3 * Vh + i;
#pragma unroll
        for (int c = 0; c < 3; ++c)
            pp[(size_t)c * Vh] = rv[c];
    }
}

template <typename S, typename VR>
void copy_r_to_p(const CgmState<S> *st, Vec<VR> r, VR *p, const size_t *off, const int *list, const int *n, int nsys,
                 hipStream_t stream)
{
    hipLaunchKernelGGL((k_copy_r_to_p<S, VR>), dim3(blas_nblk(r.Vh), nsys), dim3(kBlasNT), 0, stream, st, r, p, off,
                       list, n);
    KQ_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------ reliable updates
template <typename VR>
__global__ void __launch_bounds__(kBlasNT) k_rel_residual(const double *b, const double *ax, Vec<VR> r_lo,
                                                          const int *list, const int *n, double *part, uint32_t nblk)
{
    using V2 = c2_t<VR>;
    __shared__ double lds[kBlasNT / kWave];
    const int slot = blockIdx.y;
    if (*n <= slot)
        return;
    const int rh = list[slot];
    const uint32_t Vh = r_lo.Vh;
    const uint32_t i = blockIdx.x * kBlasNT + threadIdx.x;
    double acc = 0.0;
    if (i < Vh) {
        V2 o[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const size_t e = ((size_t)rh * 3 + c) * Vh + i;
            const double2 bb = reinterpret_cast<const double2 *>(b)[e];
            const double2 aa = reinterpret_cast<const double2 *>(ax)[e];
            const double rx = bb.x - aa.x, ry = bb.y - aa.y;
            acc += rx * rx + ry * ry;
            o[c].x = (VR)rx;
            o[c].y = (VR)ry;
        }
        r_lo.store(rh, i, o);
    }
    const double tot = block_sum<kBlasNT>(acc, lds);
    if (threadIdx.x == 0)
        part[(size_t)rh * nblk + blockIdx.x] = tot;
}

template <typename VR>
void rel_residual(const double *b, const double *ax, Vec<VR> r_lo, const int *list, const int *n, int nsys,
                  double *part, uint32_t nblk, hipStream_t stream)
{
    hipLaunchKernelGGL(k_rel_residual<VR>, dim3(nblk, nsys), dim3(kBlasNT), 0, stream, b, ax, r_lo, list, n, part,
                       nblk);
    KQ_HIP_CHECK(hipGetLastError());
}

template <typename S>
__global__ void __launch_bounds__(256) k_rel_set_rr(CgmState<S> *st, const int *list, const int *n,
                                                    const double *part, uint32_t nblk)
{
    const int slot = blockIdx.x;
    if (*n <= slot)
        return;
    const int rh = list[slot];
    const double rr = block_reduce_fixed(part + (size_t)rh * nblk, nblk);
    if (threadIdx.x == 0) {
        st[rh].rr = (S)rr;
        st[rh].r2max = rr;
        st[rh].discard = 0;
    }
}

template <typename S>
void rel_set_rr(CgmState<S> *st, const int *list, const int *n, int nsys, const double *part, uint32_t nblk,
                hipStream_t stream)
{
    hipLaunchKernelGGL(k_rel_set_rr<S>, dim3(nsys), dim3(256), 0, stream, st, list, n, part, nblk);
    KQ_HIP_CHECK(hipGetLastError());
}

template <typename S> __global__ void k_cgm_restart_list(CgmState<S> *st, const int *list, const int *n)
{
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= *n)
        return;
    CgmState<S> &c = st[list[slot]];
    c.Bm1 = (S)1.0;
    for (int k = 0; k < kMaxShift; ++k) {
        c.a[k] = (S)0.0;
        c.B[k] = (S)1.0;
        c.Z[k] = (S)1.0;
        c.Zm1[k] = (S)1.0;
    }
    c.r2max = (double)c.rr;
    c.pause = 0;
    c.hist = 0;
}

template <typename S> void cgm_restart_list(CgmState<S> *st, const int *list, const int *n, int nsys, hipStream_t stream)
{
    hipLaunchKernelGGL(k_cgm_restart_list<S>, dim3((nsys + 63) / 64), dim3(64), 0, stream, st, list, n);
    KQ_HIP_CHECK(hipGetLastError());
}

template <typename S, typename VR>
__global__ void __launch_bounds__(kBlasNT) k_rel_fold(const CgmState<S> *st, double *x_hi, VR *x_lo,
                                                      const size_t *off, uint32_t Vh, const int *list, const int *n)
{
    using V2 = c2_t<VR>;
    const int slot = blockIdx.y;
    if (*n <= slot)
        return;
    const int rh = list[slot];
    const size_t n3 = (size_t)3 * Vh;
    const size_t i = (size_t)blockIdx.x * kBlasNT + threadIdx.x;
    if (i >= n3)
        return;
    const int m = st[rh].maxTermFold;
    const int discard = st[rh].discard, live = st[rh].maxTerm;
    // streaming accesses except the store of the unshifted solution, which the FP64
    // hopping term of the reliable update reads next
    for (int k = 0; k < m; ++k) {
        const size_t e = off[rh] + (size_t)k * n3 + i;
        V2 *xl = reinterpret_cast<V2 *>(x_lo) + e;
        double2 *xh = reinterpret_cast<double2 *>(x_hi) + e;
        if (fold_keeps(discard, k, live)) {
            const V2 l = ld_nt(xl);
            double2 h = ld_nt(xh);
            h.x += (double)l.x;
            h.y += (double)l.y;
            if (k == 0)
                *xh = h;
            else
                st_nt(xh, h);
        }
        V2 z;
        z.x = 0;
        z.y = 0;
        st_nt(xl, z);
    }
}

template <typename S, typename VR>
void rel_fold(CgmState<S> *st, double *x_hi, VR *x_lo, const size_t *off, uint32_t Vh, const int *list, const int *n,
              int nsys, hipStream_t stream)
{
    const size_t n3 = (size_t)3 * Vh;
    hipLaunchKernelGGL((k_rel_fold<S, VR>), dim3((unsigned)((n3 + kBlasNT - 1) / kBlasNT), nsys), dim3(kBlasNT), 0,
                       stream, st, x_hi, x_lo, off, Vh, list, n);
    KQ_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------ deferred solution updates
// One workgroup per listed system, one thread per (shift k < maxTermFold, term m): the start
// vector (m = -1) or recorded residual m; each forms its pair of coefficients in O(J).
template <typename S>
__global__ void __launch_bounds__(256) k_seg_coef(const CgmState<S> *st, const double *shist, double *coef,
                                                  const int *list, const int *n)
{
    const int slot = blockIdx.x;
    if (*n <= slot)
        return;
    const CgmState<S> &c = st[list[slot]];
    if (c.dslot < 0)
        return;
    const double *h = shist + hist_scalar_index(c.dslot, 0, 0, 0);
    const int terms = kHistMax + 1;
    for (int idx = threadIdx.x; idx < c.maxTermFold * terms; idx += blockDim.x) {
        const int k = idx / terms, m = idx % terms - 1;
        double cx, cp;
        seg_coefficient(h, c.hist, k, m, fold_keeps(c.discard, k, c.maxTerm), cx, cp);
        double *o = coef + ((size_t)c.dslot * kMaxShift + k) * kCoefStride;
        o[1 + m] = cx;                // m = -1: o[0]
        o[kHistMax + 2 + m] = cp;     // m = -1: o[kHistMax + 1]
    }
}

template <typename S>
void seg_coef(const CgmState<S> *st, const double *shist, double *coef, const int *list, const int *n, int nsys,
              hipStream_t stream)
{
    hipLaunchKernelGGL(k_seg_coef<S>, dim3(nsys), dim3(256), 0, stream, st, shist, coef, list, n);
    KQ_HIP_CHECK(hipGetLastError());
}

// One thread per complex entry of a vector. Workgroups are dealt to the 8 XCDs round robin
// (linear id % 8), so blockIdx.x = 8 (k + kmax * (y / 8)) + y % 8 for shift k and site range y:
// the kmax workgroups of one site range run back to back on one XCD and read its recorded
// residuals once into that XCD's L2.
template <typename S, typename VR, typename XR>
__global__ void __launch_bounds__(kBlasNT) k_seg_fold(const CgmState<S> *st, const double *coef, XR *x, VR *p,
                                                      const VR *p0snap, const VR *rhist, const size_t *off,
                                                      uint32_t Vh, const int *list, const int *n, int kmax)
{
    using V2 = c2_t<VR>;
    using X2 = c2_t<XR>;
    const int slot = blockIdx.z;
    const uint32_t rest = blockIdx.x / 8;
    const int k = (int)(rest % (uint32_t)kmax);
    const size_t y = (size_t)(rest / (uint32_t)kmax) * 8 + blockIdx.x % 8;
    if (*n <= slot)
        return;
    const int rh = list[slot];
    const CgmState<S> &c = st[rh];
    if (c.dslot < 0 || k >= c.maxTermFold)
        return;
    const size_t n3 = (size_t)3 * Vh;
    const size_t e = y * kBlasNT + threadIdx.x;
    if (e >= n3)
        return;
    const bool keep_x = fold_keeps(c.discard, k, c.maxTerm);
    const bool upd_p = k >= 1 && k < c.maxTerm;
    if (!keep_x && !upd_p)
        return;
    const int J = c.hist < kHistMax ? c.hist : kHistMax;
    const double *cf = coef + ((size_t)c.dslot * kMaxShift + k) * kCoefStride;
    const size_t pe = off[rh] + (size_t)k * n3 + e;
    const V2 ps = (k == 0) ? reinterpret_cast<const V2 *>(p0snap)[(size_t)c.dslot * n3 + e]
                           : ld_nt(reinterpret_cast<const V2 *>(p) + pe);
    double xr = cf[0] * (double)ps.x, xi = cf[0] * (double)ps.y;
    double pr = cf[kHistMax + 1] * (double)ps.x, pi = cf[kHistMax + 1] * (double)ps.y;
    const V2 *rh_base = reinterpret_cast<const V2 *>(rhist) + (size_t)c.dslot * kHistMax * n3 + e;
    for (int j = 0; j < J; ++j) {
        const V2 rv = rh_base[(size_t)j * n3];
        xr += cf[1 + j] * (double)rv.x;
        xi += cf[1 + j] * (double)rv.y;
        pr += cf[kHistMax + 2 + j] * (double)rv.x;
        pi += cf[kHistMax + 2 + j] * (double)rv.y;
    }
    if (keep_x) {
        X2 *xh = reinterpret_cast<X2 *>(x) + pe;
        X2 h = ld_nt(xh);
        h.x = (XR)((double)h.x + xr);
        h.y = (XR)((double)h.y + xi);
        // the unshifted solution is read next by the FP64 hopping term of a reliable update
        if (k == 0)
            *xh = h;
        else
            st_nt(xh, h);
    }
    if (upd_p) {
        V2 np;
        np.x = (VR)pr;
        np.y = (VR)pi;
        st_nt(reinterpret_cast<V2 *>(p) + pe, np);
    }
}

template <typename S, typename VR, typename XR>
void seg_fold(const CgmState<S> *st, const double *coef, XR *x, VR *p, const VR *p0snap, const VR *rhist,
              const size_t *off, uint32_t Vh, const int *list, const int *n, int nsys, int kmax, hipStream_t stream)
{
    if (kmax <= 0)
        return;
    const size_t n3 = (size_t)3 * Vh;
    const size_t nb8 = (n3 + 8 * kBlasNT - 1) / (8 * kBlasNT);  // site ranges per XCD
    const size_t gx = 8 * (size_t)kmax * nb8;
    if (kmax > kMaxShift || gx > 0x7fffffffu)
        throw std::invalid_argument("seg_fold: grid too large");
    hipLaunchKernelGGL((k_seg_fold<S, VR, XR>), dim3((unsigned)gx, 1, nsys), dim3(kBlasNT), 0, stream, st, coef, x, p,
                       p0snap, rhist, off, Vh, list, n, kmax);
    KQ_HIP_CHECK(hipGetLastError());
}

// p0snap = p_0 and hist = 0 for the listed systems with a history slot
template <typename S, typename VR>
__global__ void __launch_bounds__(kBlasNT) k_seg_reset(CgmState<S> *st, const VR *p, VR *p0snap, const size_t *off,
                                                       uint32_t Vh, const int *list, const int *n)
{
    using V2 = c2_t<VR>;
    const int slot = blockIdx.y;
    if (*n <= slot)
        return;
    const int rh = list[slot];
    const int ds = st[rh].dslot;
    if (ds < 0)
        return;
    const size_t n3 = (size_t)3 * Vh;
    const size_t e = (size_t)blockIdx.x * kBlasNT + threadIdx.x;
    if (e < n3)
        reinterpret_cast<V2 *>(p0snap)[(size_t)ds * n3 + e] = reinterpret_cast<const V2 *>(p)[off[rh] + e];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st[rh].hist = 0;
        // with one shift left the system goes back to per-iteration updates of x_0 and p_0 (its
        // x_lo, never written while deferred, is zero): the deferral saves nothing there, and
        // x_0 then follows the FP32 p_0 the operator was applied to, as in CGM::Invert
        if (st[rh].maxTerm <= 1)
            st[rh].dslot = -1;
    }
}

template <typename S, typename VR>
void seg_reset(CgmState<S> *st, const VR *p, VR *p0snap, const size_t *off, uint32_t Vh, const int *list, const int *n,
               int nsys, hipStream_t stream)
{
    const size_t n3 = (size_t)3 * Vh;
    hipLaunchKernelGGL((k_seg_reset<S, VR>), dim3((unsigned)((n3 + kBlasNT - 1) / kBlasNT), nsys), dim3(kBlasNT), 0,
                       stream, st, p, p0snap, off, Vh, list, n);
    KQ_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------ true residuals
__global__ void __launch_bounds__(kBlasNT) k_resid2(const double2 *b, const double2 *ax, const int *list, uint32_t Vh,
                                                    double *part, uint32_t nblk)
{
    __shared__ double lds[kBlasNT / kWave];
    const int k = list ? list[blockIdx.y] : (int)blockIdx.y;
    const uint32_t i = blockIdx.x * kBlasNT + threadIdx.x;
    double acc = 0.0;
    if (i < Vh) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double2 bb = b[(size_t)c * Vh + i];
            const double2 aa = ax[((size_t)k * 3 + c) * Vh + i];
            const double rx = bb.x - aa.x, ry = bb.y - aa.y;
            acc += rx * rx + ry * ry;
        }
    }
    const double tot = block_sum<kBlasNT>(acc, lds);
    if (threadIdx.x == 0)
        part[(size_t)k * nblk + blockIdx.x] = tot;
}

void resid2_partial(const double *b, const double *ax, int nk, const int *list, uint32_t Vh, double *part,
                    uint32_t nblk, hipStream_t stream)
{
    if (nk <= 0)
        return;
    hipLaunchKernelGGL(k_resid2, dim3(nblk, nk), dim3(kBlasNT), 0, stream, reinterpret_cast<const double2 *>(b),
                       reinterpret_cast<const double2 *>(ax), list, Vh, part, nblk);
    KQ_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------ generic dot
template <typename VR>
__global__ void __launch_bounds__(kBlasNT) k_dot_partial(Vec<VR> x, Vec<VR> y, double *part, uint32_t nblk,
                                                         const int *list, const int *n)
{
    using V2 = c2_t<VR>;
    __shared__ double lds[kBlasNT / kWave];
    if (n != nullptr && *n <= (int)blockIdx.y)
        return;
    const int rh = list ? list[blockIdx.y] : (int)blockIdx.y;
    const uint32_t i = blockIdx.x * kBlasNT + threadIdx.x;
    double acc = 0.0;
    if (i < x.Vh) {
        V2 a[3], b[3];
        x.load(rh, i, a);
        y.load(rh, i, b);
#pragma unroll
        for (int c = 0; c < 3; ++c)
            acc += (double)a[c].x * (double)b[c].x + (double)a[c].y * (double)b[c].y;
    }
    const double tot = block_sum<kBlasNT>(acc, lds);
    if (threadIdx.x == 0)
        part[(size_t)rh * nblk + blockIdx.x] = tot;
}

template <typename VR>
void vec_dot_partial(Vec<VR> x, Vec<VR> y, int nrhs, double *part, uint32_t nblk, hipStream_t stream, const int *list,
                     const int *n)
{
    hipLaunchKernelGGL(k_dot_partial<VR>, dim3(nblk, nrhs), dim3(kBlasNT), 0, stream, x, y, part, nblk, list, n);
    KQ_HIP_CHECK(hipGetLastError());
}

__global__ void __launch_bounds__(256) k_reduce_partials(const double *part, uint32_t nblk, double *out)
{
    const int r = blockIdx.x;
    const double v = block_reduce_fixed(part + (size_t)r * nblk, nblk);
    if (threadIdx.x == 0)
        out[r] = v;
}

void reduce_partials(const double *part, uint32_t nblk, int nrhs, double *out, hipStream_t stream)
{
    hipLaunchKernelGGL(k_reduce_partials, dim3(nrhs), dim3(256), 0, stream, part, nblk, out);
    KQ_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------ instantiations
#define KQ_INST_S(S)                                                                                          \
    template void cgm_begin<S>(CgmState<S> *, const CgmLaunch &, hipStream_t);                                \
    template void cgm_scalars1<S>(CgmState<S> *, const CgmLaunch &, hipStream_t);                             \
    template void cgm_scalars2<S>(CgmState<S> *, const CgmLaunch &, hipStream_t);                             \
    template void cgm_init_list<S>(CgmState<S> *, const int *, const int *, int, const double *, uint32_t,    \
                                   hipStream_t);                                                              \
    template void rel_set_rr<S>(CgmState<S> *, const int *, const int *, int, const double *, uint32_t,       \
                                hipStream_t);                                                              \
    template void cgm_restart_list<S>(CgmState<S> *, const int *, const int *, int, hipStream_t);          \
    template void seg_coef<S>(const CgmState<S> *, const double *, double *, const int *, const int *, int,   \
                              hipStream_t);
KQ_INST_S(double)
KQ_INST_S(float)

#define KQ_INST_SV(S, VR)                                                                                      \
    template void cgm_axpy_norm<S, VR>(CgmState<S> *, Vec<VR>, Vec<VR>, const CgmLaunch &, hipStream_t);       \
    template void cgm_xp_update<S, VR, VR>(CgmState<S> *, VR *, VR *, Vec<VR>, const CgmLaunch &, hipStream_t); \
    template void rel_fold<S, VR>(CgmState<S> *, double *, VR *, const size_t *, uint32_t, const int *,        \
                                  const int *, int, hipStream_t);                                             \
    template void copy_r_to_p<S, VR>(const CgmState<S> *, Vec<VR>, VR *, const size_t *, const int *,          \
                                     const int *, int, hipStream_t);                                          \
    template void seg_fold<S, VR, double>(const CgmState<S> *, const double *, double *, VR *, const VR *,    \
                                          const VR *, const size_t *, uint32_t, const int *, const int *, int, int, \
                                          hipStream_t);                                                        \
    template void seg_reset<S, VR>(CgmState<S> *, const VR *, VR *, const size_t *, uint32_t, const int *,     \
                                   const int *, int, hipStream_t);
KQ_INST_SV(double, double)
KQ_INST_SV(float, float)
KQ_INST_SV(double, float)
// the float reference method folds into its FP32 solutions
template void seg_fold<float, float, float>(const CgmState<float> *, const double *, float *, float *, const float *,
                                            const float *, const size_t *, uint32_t, const int *, const int *, int, int,
                                            hipStream_t);

template void rel_residual<float>(const double *, const double *, Vec<float>, const int *, const int *, int,
                                  double *, uint32_t, hipStream_t);
template void rel_residual<double>(const double *, const double *, Vec<double>, const int *, const int *, int,
                                   double *, uint32_t, hipStream_t);
template void vec_dot_partial<double>(Vec<double>, Vec<double>, int, double *, uint32_t, hipStream_t, const int *,
                                      const int *);
template void vec_dot_partial<float>(Vec<float>, Vec<float>, int, double *, uint32_t, hipStream_t, const int *,
                                     const int *);

// ------------------------------------------------------------ small copies
__global__ void __launch_bounds__(256) k_copy_words(const unsigned long long *src, unsigned long long *dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        dst[i] = src[i];
}

void copy_words(const void *src, void *dst, size_t bytes, hipStream_t stream)
{
    if (bytes % 8 != 0)
        throw std::invalid_argument("copy_words: size must be a multiple of 8 bytes");
    const size_t n = bytes / 8;
    if (n == 0)
        return;
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 64);
    hipLaunchKernelGGL(k_copy_words, dim3(grid), dim3(256), 0, stream,
                       static_cast<const unsigned long long *>(src), static_cast<unsigned long long *>(dst), n);
    KQ_HIP_CHECK(hipGetLastError());
}

__global__ void __launch_bounds__(256) k_set_int_list(const int *src, int *list, int *count, int n)
{
    for (int i = threadIdx.x; i < n; i += 256)
        list[i] = src[i];
    if (threadIdx.x == 0)
        *count = n;
}

void set_int_list(const int *src, int *list, int *count, int n, hipStream_t stream)
{
    hipLaunchKernelGGL(k_set_int_list, dim3(1), dim3(256), 0, stream, src, list, count, n);
    KQ_HIP_CHECK(hipGetLastError());
}

}  // namespace kq
