// HISQ even-odd hopping term (fat 1-hop + Naik 3-hop) on gfx950.
//
// One thread per output site (and per block of NR right-hand sides): the
// 16 links of the site are each read once from HBM with coalesced
// structure-of-arrays loads and applied to the NR gathered neighbour vectors,
// so a batched launch reads every link once for all of its right-hand sides.
// Periodic boundaries are resolved in-kernel from the site coordinates; no
// neighbour tables are read. Follows DiracHISQ::Dslash/FwdDslash/BwdDslash
// (dirac_hisq.cpp:409-502) with the 1/2 and c3000 factors folded into the
// device links.
//
// Implementation header: kq_dslash_*.hip instantiate launch_dslash for the link
// formats (one translation unit per family, so they compile in parallel).
#pragma once

#include "kq_dslash.h"

#include <algorithm>
#include <type_traits>

namespace kq {

// gather the colour vector at site n of right-hand side j (base pointer src[j])
template <class V2>
__device__ __forceinline__ void gather(const V2 *src, uint32_t n, uint32_t Vh, V2 v[3])
{
#pragma unroll
    for (int c = 0; c < 3; ++c)
        v[c] = src[(size_t)c * Vh + n];
}

template <class A, class LK, int NR, int MU, int H>
__device__ __forceinline__ void hop_fwd(const A &a, const LK &lk, uint32_t s, const SiteCoord &sc,
                                        const typename A::V2 *const src[NR], typename A::R acc[NR][6])
{
    using R = typename A::R;
    c2_t<R> U[9];
    lk.load(a.dst_parity, MU, s, U);
    const uint32_t n = nb_index<MU>(a.g, s, sc, H);
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        typename A::V2 v[3];
        gather(src[j], n, a.g.Vh, v);
        umv_add(U, v, acc[j]);
    }
}

template <class A, class LK, int NR, int MU, int H>
__device__ __forceinline__ void hop_bwd(const A &a, const LK &lk, uint32_t s, const SiteCoord &sc,
                                        const typename A::V2 *const src[NR], typename A::R acc[NR][6])
{
    using R = typename A::R;
    const uint32_t n = nb_index<MU>(a.g, s, sc, -H);
    c2_t<R> U[9];
    lk.load(1 - a.dst_parity, MU, n, U);
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        typename A::V2 v[3];
        gather(src[j], n, a.g.Vh, v);
        udmv_sub(U, v, acc[j]);
    }
}

// Hop along x (MU 0) or y (MU 1) of a workgroup whose source rows are staged in LDS:
// st[j][c][k] holds colour c of right-hand side j at k = (y - y0 + 3) * Lx/2 + x/2 for the rows
// y0 - 3 .. y0 + NT / (Lx/2) + 2 of the workgroup's plane (yl = y - y0). Same arithmetic, in the
// same order, as hop_fwd / hop_bwd.
template <class A, class LK, int NR, int MU, int H, bool BWD, int KR>
__device__ __forceinline__ void hop_staged(const A &a, const LK &lk, uint32_t s, const SiteCoord &sc, int yl,
                                           const typename A::V2 (*st)[3][KR], typename A::R acc[NR][6])
{
    using R = typename A::R;
    constexpr int XH = kStageLxh;
    const int h = BWD ? -H : H;
    const int k = MU == 0 ? (yl + 3) * XH + (wrap_add(sc.c[0], h, a.g.L[0]) >> 1) : (yl + 3 + h) * XH + (sc.c[0] >> 1);
    c2_t<R> U[9];
    if (BWD)
        lk.load(1 - a.dst_parity, MU, nb_index<MU>(a.g, s, sc, h), U);
    else
        lk.load(a.dst_parity, MU, s, U);
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        typename A::V2 v[3];
#pragma unroll
        for (int c = 0; c < 3; ++c)
            v[c] = st[j][c][k];
        if (BWD)
            udmv_sub(U, v, acc[j]);
        else
            umv_add(U, v, acc[j]);
    }
}

// The staged kernel's 16 hops as one chain in which each hop's link words are requested before
// the previous hop's arithmetic (a one-hop prefetch; hop I: 0-3 x/y forward from LDS, 4-7 z/t
// forward from global memory, 8-11 and 12-15 the same backward; even I fat, odd I long links).
// A scheduling fence after each prefetch keeps the compiler from sinking it back. Measured at 32^4
// with FP32 links: 4 right-hand sides -2 %, 8 -8 %, 2 and 3 within 1 % (so used from 4 on, FP32
// links only: the FP64 link words would take 28-36 more registers).
template <int I> struct StagedHop {
    static constexpr int kMu = (I % 8) / 2;  // 0, 1 staged (x, y); 2, 3 global (z, t)
    static constexpr bool kBwd = I >= 8, kLong = (I % 2) == 1;
    static constexpr int kH = kLong ? 3 : 1;
};

template <class A, int I>
__device__ __forceinline__ auto staged_fetch(const A &a, uint32_t s, const SiteCoord &sc, uint32_t &n)
{
    using H = StagedHop<I>;
    n = nb_index<H::kMu>(a.g, s, sc, H::kBwd ? -H::kH : H::kH);
    const uint32_t ls = H::kBwd ? n : s;
    const int q = H::kBwd ? 1 - a.dst_parity : a.dst_parity;
    if constexpr (H::kLong)
        return a.lng.fetch(q, H::kMu, ls);
    else
        return a.fat.fetch(q, H::kMu, ls);
}

template <class A, int NR, int KR, int I, class RW>
__device__ __forceinline__ void staged_chain(const A &a, uint32_t s, const SiteCoord &sc, int yl,
                                             const typename A::V2 (*st)[3][KR], const typename A::V2 *const src[NR],
                                             typename A::R acc[NR][6], const RW &cur, uint32_t n)
{
    using H = StagedHop<I>;
    constexpr int XH = kStageLxh;
    typename A::V2 v[NR][3];
    if constexpr (H::kMu < 2) {
        const int h = H::kBwd ? -H::kH : H::kH;
        const int k = H::kMu == 0 ? (yl + 3) * XH + (wrap_add(sc.c[0], h, a.g.L[0]) >> 1)
                                  : (yl + 3 + h) * XH + (sc.c[0] >> 1);
#pragma unroll
        for (int j = 0; j < NR; ++j)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                v[j][c] = st[j][c][k];
    } else {
#pragma unroll
        for (int j = 0; j < NR; ++j)
            gather(src[j], n, a.g.Vh, v[j]);
    }
    uint32_t n2 = 0;
    if constexpr (I < 15) {
        const auto nxt = staged_fetch<A, I + 1>(a, s, sc, n2);
        __builtin_amdgcn_sched_barrier(0);
        c2_t<typename A::R> U[9];
        if constexpr (H::kLong)
            a.lng.unpack(cur, U);
        else
            a.fat.unpack(cur, U);
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            if constexpr (H::kBwd)
                udmv_sub(U, v[j], acc[j]);
            else
                umv_add(U, v[j], acc[j]);
        }
        staged_chain<A, NR, KR, I + 1>(a, s, sc, yl, st, src, acc, nxt, n2);
    } else {
        c2_t<typename A::R> U[9];
        if constexpr (H::kLong)
            a.lng.unpack(cur, U);
        else
            a.fat.unpack(cur, U);
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            if constexpr (H::kBwd)
                udmv_sub(U, v[j], acc[j]);
            else
                umv_add(U, v[j], acc[j]);
        }
    }
}

// output store: streaming for the FP64 and FP32 link formats (measured 5% and 4%
// faster at 32^4), plain for the 16-bit one (3% slower with streaming stores). Also
// streaming where the next hopping term gathers the output (MdaggM's intermediate):
// plain stores there made the 32^4 solves 1-2% slower. The choice must stay
// compile-time: a run-time flag in this store made the FP64 term almost 2x slower.
template <class LF, class V2> __device__ __forceinline__ void put(V2 *p, V2 v)
{
    if constexpr (std::is_same<LF, LinkH>::value)
        *p = v;
    else
        st_nt(p, v);
}

// Epilogue of the hopping term: store D src (EPI_STORE), or dst = sigma aux - D src with the
// per-workgroup partial sums of <aux, dst> (EPI_AXMY_DOT)
template <class LF, typename SR, int NR, int EPI, int NT, class A, typename R>
__device__ __forceinline__ void dslash_epilogue(const A &a, const R (&acc)[NR][6], c2_t<SR> *const (&dst)[NR],
                                                const c2_t<SR> *const (&aux)[NR], const int (&rr)[NR], int j0,
                                                int nact, uint32_t s, bool valid, double *lds)
{
    using V2 = c2_t<SR>;
    const uint32_t Vh = a.g.Vh;
    if constexpr (EPI == EPI_STORE) {
        if (valid) {
#pragma unroll
            for (int j = 0; j < NR; ++j) {
                // padding slots repeat the last live right-hand side, so the
                // store is unconditional (same value to the same address)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    put<LF>(&dst[j][(size_t)c * Vh + s], mk<V2>((SR)acc[j][2 * c], (SR)acc[j][2 * c + 1]));
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            const bool live = j0 + j < nact;  // uniform over the block
            double d = 0.0;
            if (valid && live) {
                const double sg = a.sigma[rr[j]];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const V2 p = aux[j][(size_t)c * Vh + s];
                    const SR ox = (SR)((R)sg * (R)p.x - acc[j][2 * c]);
                    const SR oy = (SR)((R)sg * (R)p.y - acc[j][2 * c + 1]);
                    put<LF>(&dst[j][(size_t)c * Vh + s], mk<V2>(ox, oy));
                    d += (double)p.x * (double)ox + (double)p.y * (double)oy;
                }
            }
            const double tot = block_sum<NT>(d, lds);
            if (threadIdx.x == 0 && live)
                a.partial[(size_t)rr[j] * a.nblk + blockIdx.x] = tot;
            __syncthreads();
        }
    }
}

// Per right-hand-side base pointers of a workgroup's block of NR systems, resolved once (uniform
// over the block); padding slots beyond the nact live ones repeat the last live system
template <int NR, class A>
__device__ __forceinline__ void dslash_rhs(const A &a, int j0, int nact, int (&rr)[NR], typename A::V2 *(&dst)[NR],
                                           const typename A::V2 *(&src)[NR], const typename A::V2 *(&aux)[NR], bool with_aux)
{
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        const int jj = (j0 + j < nact) ? j0 + j : nact - 1;
        rr[j] = a.rlist ? a.rlist[jj] : jj;
        src[j] = a.src.p + a.src.base(rr[j]);
        dst[j] = a.dst.p + a.dst.base(rr[j]);
        aux[j] = with_aux ? a.aux.p + a.aux.base(rr[j]) : nullptr;
    }
}

template <class LF, class LL, typename SR, int NR, int EPI, int NT>
__global__ void __launch_bounds__(NT) k_dslash(const DslashArgs<LF, LL, SR> a)
{
    using A = DslashArgs<LF, LL, SR>;
    using R = typename A::R;
    using V2 = c2_t<SR>;
    __shared__ double lds[NT / kWave];
    if (a.run != nullptr && *a.run == 0)
        return;  // uniform over the grid
    const int nact = a.nact ? *a.nact : a.nrhs;
    const int j0 = blockIdx.y * NR;
    if (j0 >= nact)
        return;  // uniform over the block
    int rr[NR];
    const V2 *src[NR];
    V2 *dst[NR];
    const V2 *aux[NR];
    dslash_rhs<NR>(a, j0, nact, rr, dst, src, aux, EPI == EPI_AXMY_DOT);
    bool valid;
    const uint32_t s = tile_site<NT>(a.g, blockIdx.x, gridDim.x, threadIdx.x, valid);
    R acc[NR][6];
#pragma unroll
    for (int j = 0; j < NR; ++j)
#pragma unroll
        for (int k = 0; k < 6; ++k)
            acc[j][k] = 0;
    if (valid) {
        const SiteCoord sc = site_coords(a.g, s, a.dst_parity);
        hop_fwd<A, LF, NR, 0, 1>(a, a.fat, s, sc, src, acc);
        hop_fwd<A, LL, NR, 0, 3>(a, a.lng, s, sc, src, acc);
        hop_fwd<A, LF, NR, 1, 1>(a, a.fat, s, sc, src, acc);
        hop_fwd<A, LL, NR, 1, 3>(a, a.lng, s, sc, src, acc);
        hop_fwd<A, LF, NR, 2, 1>(a, a.fat, s, sc, src, acc);
        hop_fwd<A, LL, NR, 2, 3>(a, a.lng, s, sc, src, acc);
        hop_fwd<A, LF, NR, 3, 1>(a, a.fat, s, sc, src, acc);
        hop_fwd<A, LL, NR, 3, 3>(a, a.lng, s, sc, src, acc);
        hop_bwd<A, LF, NR, 0, 1>(a, a.fat, s, sc, src, acc);
        hop_bwd<A, LL, NR, 0, 3>(a, a.lng, s, sc, src, acc);
        hop_bwd<A, LF, NR, 1, 1>(a, a.fat, s, sc, src, acc);
        hop_bwd<A, LL, NR, 1, 3>(a, a.lng, s, sc, src, acc);
        hop_bwd<A, LF, NR, 2, 1>(a, a.fat, s, sc, src, acc);
        hop_bwd<A, LL, NR, 2, 3>(a, a.lng, s, sc, src, acc);
        hop_bwd<A, LF, NR, 3, 1>(a, a.fat, s, sc, src, acc);
        hop_bwd<A, LL, NR, 3, 3>(a, a.lng, s, sc, src, acc);
    }
    dslash_epilogue<LF, SR, NR, EPI, NT>(a, acc, dst, aux, rr, j0, nact, s, valid, lds);
}

// The batched hopping term with the x and y neighbours read from LDS: the workgroup (256 sites,
// whole x-rows of one (x, y) plane, stage_rows()) first copies its rows and their +-3
// y-neighbour rows of every source into LDS; the z and t neighbours come from global memory.
// Used for NR >= 2, where every staged row serves several right-hand sides' gathers: measured at
// 32^4 with FP32 links, 4 right-hand sides 130 -> 116 us, 2 and 8 about 4 % faster (1 would be
// a little slower). Results are bitwise those of k_dslash.
template <class LF, class LL, typename SR, int NR, int EPI>
__global__ void __launch_bounds__(256) k_dslash_staged(const DslashArgs<LF, LL, SR> a)
{
    using A = DslashArgs<LF, LL, SR>;
    using R = typename A::R;
    using V2 = c2_t<SR>;
    constexpr int NT = 256, XH = kStageLxh, KR = NT + 6 * XH;
    __shared__ __align__(16) V2 st[NR][3][KR];
    __shared__ double lds[NT / kWave];
    if (a.run != nullptr && *a.run == 0)
        return;  // uniform over the grid
    const int nact = a.nact ? *a.nact : a.nrhs;
    const int j0 = blockIdx.y * NR;
    if (j0 >= nact)
        return;  // uniform over the block
    int rr[NR];
    const V2 *src[NR];
    V2 *dst[NR];
    const V2 *aux[NR];
    dslash_rhs<NR>(a, j0, nact, rr, dst, src, aux, EPI == EPI_AXMY_DOT);
    const uint32_t Vh = a.g.Vh;
    // true for every thread where this kernel is launched (tile order); the hops and the epilogue
    // stay under it as in k_dslash: with the hops outside, the compiler moved them into the
    // epilogue's branches and used twice the registers, with spills
    bool valid;
    const uint32_t s = tile_site<NT>(a.g, blockIdx.x, gridDim.x, threadIdx.x, valid);
    const SiteCoord sc = site_coords(a.g, s, a.dst_parity);
    const int yl = (int)threadIdx.x / XH;
    const int y0 = sc.c[1] - yl;  // uniform over the workgroup
    const uint32_t pbase = s - threadIdx.x - (uint32_t)(XH * y0);
    const int L1 = a.g.L[1];
    // rows are XH entries of one colour plane, so two FP32 entries (16 bytes) load together when
    // every source is 16-byte aligned (Vh even: so are its colour planes and rows)
    bool wide = sizeof(V2) == 8 && (Vh % 2u) == 0;
#pragma unroll
    for (int j = 0; j < NR; ++j)
        wide = wide && ((reinterpret_cast<uintptr_t>(src[j]) & 15u) == 0);
    if (wide) {
        constexpr int KH = KR / 2;
        for (int e = threadIdx.x; e < NR * 3 * KH; e += NT) {
            const int jc = e / KH, kk = 2 * (e - jc * KH), j = jc / 3, c = jc - 3 * j;
            const int row = kk / XH, xh = kk - row * XH;
            int ys = y0 - 3 + row;
            ys = ys < 0 ? ys + L1 : (ys >= L1 ? ys - L1 : ys);
            const uint32_t n = pbase + (uint32_t)(XH * ys + xh);
            *reinterpret_cast<kq_v4f *>(&st[j][c][kk]) =
                *reinterpret_cast<const kq_v4f *>(a.src.p + a.src.base(rr[j]) + (size_t)c * Vh + n);
        }
    } else {
        for (int e = threadIdx.x; e < NR * 3 * KR; e += NT) {
            const int jc = e / KR, kk = e - jc * KR, j = jc / 3, c = jc - 3 * j;
            const int row = kk / XH, xh = kk - row * XH;
            int ys = y0 - 3 + row;
            ys = ys < 0 ? ys + L1 : (ys >= L1 ? ys - L1 : ys);
            const uint32_t n = pbase + (uint32_t)(XH * ys + xh);
            st[j][c][kk] = a.src.p[a.src.base(rr[j]) + (size_t)c * Vh + n];
        }
    }
    __syncthreads();
    R acc[NR][6];
#pragma unroll
    for (int j = 0; j < NR; ++j)
#pragma unroll
        for (int k = 0; k < 6; ++k)
            acc[j][k] = 0;
    if (valid) {
        if constexpr (NR >= 4 && std::is_same<R, float>::value) {
            uint32_t n0;
            const auto r0 = staged_fetch<A, 0>(a, s, sc, n0);
            staged_chain<A, NR, KR, 0>(a, s, sc, yl, st, src, acc, r0, n0);
        } else {
            hop_staged<A, LF, NR, 0, 1, false, KR>(a, a.fat, s, sc, yl, st, acc);
            hop_staged<A, LL, NR, 0, 3, false, KR>(a, a.lng, s, sc, yl, st, acc);
            hop_staged<A, LF, NR, 1, 1, false, KR>(a, a.fat, s, sc, yl, st, acc);
            hop_staged<A, LL, NR, 1, 3, false, KR>(a, a.lng, s, sc, yl, st, acc);
            hop_fwd<A, LF, NR, 2, 1>(a, a.fat, s, sc, src, acc);
            hop_fwd<A, LL, NR, 2, 3>(a, a.lng, s, sc, src, acc);
            hop_fwd<A, LF, NR, 3, 1>(a, a.fat, s, sc, src, acc);
            hop_fwd<A, LL, NR, 3, 3>(a, a.lng, s, sc, src, acc);
            hop_staged<A, LF, NR, 0, 1, true, KR>(a, a.fat, s, sc, yl, st, acc);
            hop_staged<A, LL, NR, 0, 3, true, KR>(a, a.lng, s, sc, yl, st, acc);
            hop_staged<A, LF, NR, 1, 1, true, KR>(a, a.fat, s, sc, yl, st, acc);
            hop_staged<A, LL, NR, 1, 3, true, KR>(a, a.lng, s, sc, yl, st, acc);
            hop_bwd<A, LF, NR, 2, 1>(a, a.fat, s, sc, src, acc);
            hop_bwd<A, LL, NR, 2, 3>(a, a.lng, s, sc, src, acc);
            hop_bwd<A, LF, NR, 3, 1>(a, a.fat, s, sc, src, acc);
            hop_bwd<A, LL, NR, 3, 3>(a, a.lng, s, sc, src, acc);
        }
    }
    dslash_epilogue<LF, SR, NR, EPI, NT>(a, acc, dst, aux, rr, j0, nact, s, valid, lds);
}

// LDS of a workgroup the device allows
static int lds_block_bytes()
{
    static const int lds_block = [] {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
            v = 65536;
        return v;
    }();
    return lds_block;
}

// dynamic LDS bytes that let at most `cap` waves per SIMD (4 SIMDs per CU) be resident
static size_t cap_lds_bytes(int cap, int nt)
{
    if (cap <= 0)
        return 0;
    static const int lds_cu = [] {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess || v <= 0)
            v = 65536;
        return v;
    }();
    const int lds_block = lds_block_bytes();
    const int blocks = std::max(1, cap * 4 / std::max(1, nt / kWave));
    // a little under lds_cu / blocks, so that `blocks` fit and blocks + 1 do not
    const int bytes = lds_cu / blocks - 256;
    return (size_t)std::max(0, std::min(bytes, lds_block - 256));
}

template <class LF, class LL, typename SR, int NR, int EPI>
static void launch_nt(const DslashArgs<LF, LL, SR> &a, int nt, int cap, hipStream_t st)
{
    const int nact_max = a.nrhs;
    dim3 grid((a.g.Vh + nt - 1) / nt, (nact_max + NR - 1) / NR);
    const size_t lds = cap_lds_bytes(cap, nt);
    if constexpr (NR >= 2) {
        // staged rows (static LDS of the kernel): NR * 3 colour entries for each of the
        // nt + 6 Lx/2 sites of the workgroup's rows and their +-3 y-neighbour rows; an occupancy
        // cap adds dynamic LDS only beyond that
        const size_t sb = (size_t)NR * 3 * (256 + 6 * kStageLxh) * sizeof(c2_t<SR>);
        if (dslash_staged(a.g, NR, nt) && sb + 1024 <= (size_t)lds_block_bytes()) {
            hipLaunchKernelGGL((k_dslash_staged<LF, LL, SR, NR, EPI>), grid, dim3(256), lds > sb ? lds - sb : 0,
                               st, a);
            KQ_HIP_CHECK(hipGetLastError());
            return;
        }
    }
    switch (nt) {
    case 64:
        hipLaunchKernelGGL((k_dslash<LF, LL, SR, NR, EPI, 64>), grid, dim3(64), lds, st, a);
        break;
    case 128:
        hipLaunchKernelGGL((k_dslash<LF, LL, SR, NR, EPI, 128>), grid, dim3(128), lds, st, a);
        break;
    case 256:
        hipLaunchKernelGGL((k_dslash<LF, LL, SR, NR, EPI, 256>), grid, dim3(256), lds, st, a);
        break;
    default:
        throw std::runtime_error("dslash: unsupported block size " + std::to_string(nt));
    }
    KQ_HIP_CHECK(hipGetLastError());
}

template <class LF, class LL, typename SR, int NR>
static void launch_epi(const DslashArgs<LF, LL, SR> &a, int epi, int nt, int cap, hipStream_t st)
{
    if (epi == EPI_STORE)
        launch_nt<LF, LL, SR, NR, EPI_STORE>(a, nt, cap, st);
    else
        launch_nt<LF, LL, SR, NR, EPI_AXMY_DOT>(a, nt, cap, st);
}

template <class LF, class LL, typename SR>
void launch_dslash(DslashArgs<LF, LL, SR> a, int epi, int nr, int nt, hipStream_t st, int cap)
{
    if (a.g.Vh == 0 || a.nrhs <= 0)
        return;
    a.nblk = (a.g.Vh + nt - 1) / nt;
    switch (nr) {
    case 1:
        launch_epi<LF, LL, SR, 1>(a, epi, nt, cap, st);
        break;
    case 2:
        launch_epi<LF, LL, SR, 2>(a, epi, nt, cap, st);
        break;
    case 3:
        launch_epi<LF, LL, SR, 3>(a, epi, nt, cap, st);
        break;
    case 4:
        launch_epi<LF, LL, SR, 4>(a, epi, nt, cap, st);
        break;
    case 8:
        launch_epi<LF, LL, SR, 8>(a, epi, nt, cap, st);
        break;
    default:
        throw std::runtime_error("dslash: unsupported RHS blocking " + std::to_string(nr));
    }
}


}  // namespace kq
