// Solve drivers: the reference multishift CG in the loaded precision, and the
// FP64 / mixed-precision multishift CG with per-shift FP64 verification and
// batched refinement.
#include "kq_solve.h"
#include "kq_blas.h"
#include "kq_cgm_host.h"
#include "kq_learned.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace kq {

using clk = std::chrono::steady_clock;
static double secs(clk::time_point a, clk::time_point b)
{
    return std::chrono::duration<double>(b - a).count();
}

std::vector<double> reduce_host(const double *part_dev, uint32_t nblk, int n, hipStream_t st)
{
    DevBuf out(sizeof(double) * n);
    reduce_partials(part_dev, nblk, n, out.as<double>(), st);
    std::vector<double> h(n);
    KQ_HIP_CHECK(hipMemcpyAsync(h.data(), out.as<void>(), sizeof(double) * n, hipMemcpyDeviceToHost, st));
    KQ_HIP_CHECK(hipStreamSynchronize(st));
    return h;
}

static void print_iters(int iter, int max_iter)
{
    // inverter_cgm.cpp:80-84
    if (iter >= max_iter - 1)
        log_info("CGM: Warning max iteration reached %d\n", iter);
    else
        log_info("CGM: # iterations %d\n", iter);
}

// occupancy caps of the hopping term for every blocking the runner may use, and the
// FP64 term of the reliable updates (block size and caps)
// KQ_CAP=<waves per SIMD> (0..8) replaces every autotuned occupancy cap of the solvers' hopping
// terms (a tuning switch: results are bitwise the same, only the resident waves differ)
static int cap_override()
{
    static const int v = [] {
        const char *e = getenv("KQ_CAP");
        if (e == nullptr)
            return -1;
        const int n = atoi(e);
        return (n >= 0 && n <= 8) ? n : -1;
    }();
    return v;
}

static void set_caps(RunnerTuning &t, int slot, int mode, bool reliable)
{
    const bool dbl = mode == KQ_MODE_FP64 || (mode == KQ_MODE_REFERENCE && ctx().fat[slot].prec == KUQCD_DOUBLE);
    static const int kNrs[5] = {1, 2, 3, 4, 8};
    for (int i = 0; i < 5; ++i)
        t.cap[i] = (dbl && kNrs[i] > 2) ? 0 : tuned_cap(slot, mode, kNrs[i], t.nt);
    const int ov = cap_override();
    if (ov >= 0)
        for (int i = 0; i < 5; ++i)
            t.cap[i] = ov;
    if (reliable) {
        t.nt64 = tuned_launch(slot, KQ_MODE_FP64, 1).second;
        // KQ_RELIABLE_NT=64|128|256 fixes the FP64 term's block size (to test the runner
        // with block sizes that differ between its two hopping terms)
        if (const char *e = getenv("KQ_RELIABLE_NT")) {
            const int v = atoi(e);
            if (v == 64 || v == 128 || v == 256)
                t.nt64 = v;
        }
        t.cap64[0] = tuned_cap(slot, KQ_MODE_FP64, 1, t.nt64);
        t.cap64[1] = tuned_cap(slot, KQ_MODE_FP64, 2, t.nt64);
        if (ov >= 0)
            t.cap64[0] = t.cap64[1] = ov;
    }
}

// ---------------------------------------------------------------- reference method
template <typename T, class LF, class LL>
static KuQCDStatus solve_ref_t(int slot, LF fat, LL lng, void **x, void *b, KuQCDInvertParam &p)
{
    Context &c = ctx();
    const Geom g = slot_geom(slot);
    const int parity = (p.matpc_type == KUQCD_EVEN_EVEN) ? 0 : 1;
    const int ns = p.num_shift;
    RunnerTuning tune;
    tune.nr = 1;
    tune.nt = kReproNT;
    set_caps(tune, slot, KQ_MODE_REFERENCE, false);
    tune.chunk = 16;
    CgmRunner<T, T, LF, LL> run(g, parity, fat, lng, 1, ns, c.stream, tune);
    // the solutions take their updates per segment of recorded residuals: the same iterations
    // and scalars as updating every shift each iteration, with a quarter of the vector traffic
    run.enable_deferred(deferred_slots<T>(1, g.Vh));
    DevBuf stage, bdev(sizeof(c2_t<T>) * 3 * (size_t)g.Vh);
    const size_t n3 = (size_t)3 * g.Vh;
    upload_vec<T, T>(b, Vec<T>{bdev.as<c2_t<T>>(), g.Vh, n3}, 0, stage, c.stream);
    std::vector<T> sh(ns);
    for (int k = 0; k < ns; ++k)
        sh[k] = (T)p.shifts[k];  // kuqcd.cpp:270-272
    const std::vector<double> tolk(ns, p.tol);
    run.add_system(ns, sh.data(), tolk.data(), p.max_iter, bdev.as<void>(), nullptr);
    // A shift the solver drops is final once the fold enqueued with its drop has run: its copy
    // to the caller's array goes out on the copy stream while the next chunk of iterations runs
    // (the host blocks in that copy, not the GPU), so the transfers of all but the last shifts
    // overlap the solve.
    std::vector<char> copied(ns, 0);
    std::vector<std::pair<int, hipEvent_t>> ready;
    struct EventGuard {  // events of drops not yet copied when an error ends the solve
        std::vector<std::pair<int, hipEvent_t>> &v;
        ~EventGuard()
        {
            for (const auto &kv : v)
                (void)hipEventDestroy(kv.second);
        }
    } event_guard{ready};
    DevBuf stage2(sizeof(T) * 6 * (size_t)g.Vh);
    auto on_drop = [&](int, int k) {
        hipEvent_t e;
        KQ_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        KQ_HIP_CHECK(hipEventRecord(e, c.stream));
        ready.emplace_back(k, e);
    };
    auto after_launch = [&] {
        for (const auto &kv : ready) {
            KQ_HIP_CHECK(hipStreamWaitEvent(c.copy_stream, kv.second, 0));
            Vec<T> xv{reinterpret_cast<c2_t<T> *>(run.x_lo(0, kv.first)), g.Vh, n3};
            download_vec<T, T>(xv, 0, x[kv.first], stage2, c.copy_stream);
            (void)hipEventDestroy(kv.second);
            copied[kv.first] = 1;
        }
        ready.clear();
    };
    run.run(on_drop, after_launch);
    after_launch();
    const int iter = run.state(0).iter;
    for (int k = 0; k < ns; ++k) {
        if (copied[k])
            continue;
        Vec<T> xv{reinterpret_cast<c2_t<T> *>(run.x_lo(0, k)), g.Vh, n3};
        download_vec<T, T>(xv, 0, x[k], stage, c.stream);
    }
    p.iter = iter;
    print_iters(iter, p.max_iter);
    return KUQCD_SUCCESS;
}

// ---------------------------------------------------------------- full operator
// KUQCD_HISQ_DIRAC. The reference's kuqcd_inverter allocates the solve's fields on the parity of
// matpc_type (kuqcd.cpp:200, 214-252) while the non-PC operator addresses both parities
// (dirac_hisq.cpp:130, 422-428), so it reads and writes out of bounds. Here b and every x_k are
// fields of both parities: V sites in the host order of a full field, the even block and then
// the odd one (kuqcd.cpp:129-169 with i in [0, V)). CGM::Invert (inverter_cgm.cpp:17-87) runs
// on them with A_k = sigma_k - D D, D the hopping term over all sites (dirac_hisq.cpp:409-462
// with offseti = offsetg = 0): D maps the odd block to the even one and the even to the odd, so
// D D acts on each parity by itself. The recurrence runs on the host in T (float for
// KUQCD_SINGLE, as the reference's scalars); hopping terms and vector updates run on the device,
// dot products as FP64 partial sums rounded to T. matpc_type is checked as the reference checks
// it and not used otherwise.
template <typename T, class LF, class LL> struct FullOps {
    using V2 = c2_t<T>;
    Geom g;
    LF fat;
    LL lng;
    hipStream_t st;
    size_t n3, n;  // complex entries of one parity's field, of a field of both parities
    V2 *r, *ap, *tmp, *p, *x;
    double *part, *out;
    double host_out = 0.0;

    V2 *pk(int k) const { return p + (size_t)k * n; }
    V2 *xk(int k) const { return x + (size_t)k * n; }
    // dst = D src over both parities: dst(q) = D src(1 - q)
    void dslash(const V2 *src, V2 *dst) const
    {
        for (int q = 0; q < 2; ++q) {
            DslashArgs<LF, LL, T> a{};
            a.g = g;
            a.dst_parity = q;
            a.fat = fat;
            a.lng = lng;
            a.src = Vec<T>{const_cast<V2 *>(src) + (size_t)(1 - q) * n3, g.Vh, n3};
            a.dst = Vec<T>{dst + (size_t)q * n3, g.Vh, n3};
            a.nrhs = 1;
            launch_dslash(a, EPI_STORE, 1, kReproNT, st, 0);
        }
    }
    T dot(const V2 *u, const V2 *v)
    {
        flat_dot<T>(u, v, n, part, out, st);
        KQ_HIP_CHECK(hipMemcpyAsync(&host_out, out, sizeof(double), hipMemcpyDeviceToHost, st));
        KQ_HIP_CHECK(hipStreamSynchronize(st));
        return (T)host_out;
    }
    T norm2_r() { return dot(r, r); }
    T apply(T sigma0)
    {
        dslash(pk(0), tmp);
        dslash(tmp, ap);
        flat_axpby<T>(pk(0), ap, n, sigma0, T(-1), st);  // Axmy: ap = sigma_0 p0 - D D p0
        return dot(pk(0), ap);
    }
    void update_r(T b0) { flat_axpby<T>(ap, r, n, b0, T(1), st); }
    void update_x(int k, T bk) { flat_axpby<T>(pk(k), xk(k), n, -bk, T(1), st); }
    void update_p(int k, T zk, T ak) { flat_axpby<T>(r, pk(k), n, zk, ak, st); }
};

template <typename T, class LF, class LL>
static KuQCDStatus solve_full_t(int slot, LF fat, LL lng, void **x, void *b, KuQCDInvertParam &p)
{
    using V2 = c2_t<T>;
    Context &c = ctx();
    const Geom g = slot_geom(slot);
    const int ns = p.num_shift;
    const size_t n3 = (size_t)3 * g.Vh, n = 2 * n3;
    DevBuf bufs(sizeof(V2) * n * (3 + 2 * (size_t)ns)), part(sizeof(double) * (flat_dot_nblk(n) + 1)), stage;
    V2 *base = bufs.as<V2>();
    FullOps<T, LF, LL> ops{g, fat, lng, c.stream, n3, n, base, base + n, base + 2 * n, base + 3 * n,
                       base + (3 + (size_t)ns) * n, part.as<double>(), part.as<double>() + flat_dot_nblk(n)};
    const Vec<T> rv{ops.r, g.Vh, n3};
    for (int q = 0; q < 2; ++q)
        upload_vec<T, T>(static_cast<const T *>(b) + (size_t)6 * g.Vh * q, rv, q, stage, c.stream);
    for (int k = 0; k < ns; ++k)
        KQ_HIP_CHECK(hipMemcpyAsync(ops.pk(k), ops.r, sizeof(V2) * n, hipMemcpyDeviceToDevice, c.stream));
    KQ_HIP_CHECK(hipMemsetAsync(ops.x, 0, sizeof(V2) * n * ns, c.stream));
    std::vector<T> sh(ns);
    for (int k = 0; k < ns; ++k)
        sh[k] = (T)p.shifts[k];  // kuqcd.cpp:270-272
    const int iter = cgm_recurrence<T>(ops, ns, sh.data(), p.tol, p.max_iter);
    for (int k = 0; k < ns; ++k)
        for (int q = 0; q < 2; ++q)
            download_vec<T, T>(Vec<T>{ops.xk(k), g.Vh, n3}, q, static_cast<T *>(x[k]) + (size_t)6 * g.Vh * q, stage,
                               c.stream);
    p.iter = iter;
    print_iters(iter, p.max_iter);
    return KUQCD_SUCCESS;
}

KuQCDStatus solve_reference(int slot, void **x, void *b, KuQCDInvertParam &p)
{
    Context &c = ctx();
    const bool full = p.dslash_type == KUQCD_HISQ_DIRAC;
    if (c.fat[slot].prec == KUQCD_DOUBLE) {
        KuQCDStatus st = KUQCD_SUCCESS;
        with_fp64_links(slot, [&](auto fat, auto lng) {
            using LF = decltype(fat);
            using LL = decltype(lng);
            st = full ? solve_full_t<double, LF, LL>(slot, fat, lng, x, b, p)
                      : solve_ref_t<double, LF, LL>(slot, fat, lng, x, b, p);
        });
        return st;
    }
    KuQCDStatus st = KUQCD_SUCCESS;
    with_fp32_links(slot, [&](auto fat, auto lng) {
        using LF = decltype(fat);
        using LL = decltype(lng);
        st = full ? solve_full_t<float, LF, LL>(slot, fat, lng, x, b, p) : solve_ref_t<float, LF, LL>(slot, fat, lng, x, b, p);
    });
    return st;
}

// ---------------------------------------------------------------- FP64 verification
// res2[k] = |b - A_k x_k|^2 for the k in `which` (all nk when empty), x[k] FP64
// SoA with stride 3Vh, A_k = sigma_k - D D with absolute shifts sigma_k. The
// FP64 hopping term runs batched over the listed vectors (links read once per
// pair of vectors).
template <class LL>
static std::vector<double> true_res2_t(int slot, LinkD fat, LL lng, int parity, const double *b, const double *x,
                                       int nk, const std::vector<double> &sigma_abs, std::vector<int> which)
{
    Context &c = ctx();
    const Geom g = slot_geom(slot);
    const size_t n3 = (size_t)3 * g.Vh;
    if (which.empty())
        for (int k = 0; k < nk; ++k)
            which.push_back(k);
    const int nl = (int)which.size();
    DevBuf tmp(sizeof(double2) * nk * n3), ax(sizeof(double2) * nk * n3), sig(sizeof(double) * nk);
    DevBuf list(sizeof(int) * (nl + 1));
    // FP64 blocking and occupancy cap from the autotuner (one group of at most two
    // right-hand sides per workgroup); the fixed block size keeps the residuals reproducible
    const auto pick = tuned_launch(slot, KQ_MODE_FP64, nl);
    const int nr = std::min(pick.first, 2), nt = kReproNT;
    const int cap = tuned_cap(slot, KQ_MODE_FP64, nr, nt);
    const uint32_t nblk_d = dslash_nblk(g.Vh, nt);
    DevBuf part(sizeof(double) * nk * std::max(nblk_d, blas_nblk(g.Vh)));
    std::vector<int> hl(which);
    hl.push_back(nl);
    KQ_HIP_CHECK(hipMemcpyAsync(sig.as<void>(), sigma_abs.data(), sizeof(double) * nk, hipMemcpyHostToDevice,
                                c.stream));
    KQ_HIP_CHECK(hipMemcpyAsync(list.as<void>(), hl.data(), sizeof(int) * (nl + 1), hipMemcpyHostToDevice, c.stream));
    DslashArgs<LinkD, LL, double> a{};
    a.g = g;
    a.fat = fat;
    a.lng = lng;
    a.nrhs = nl;
    a.rlist = list.as<int>();
    a.nact = list.as<int>() + nl;
    Vec<double> xv{reinterpret_cast<double2 *>(const_cast<double *>(x)), g.Vh, n3};
    Vec<double> tv{tmp.as<double2>(), g.Vh, n3}, av{ax.as<double2>(), g.Vh, n3};
    a.dst_parity = 1 - parity;
    a.src = xv;
    a.dst = tv;
    launch_dslash(a, EPI_STORE, nr, nt, c.stream, cap);
    a.dst_parity = parity;
    a.src = tv;
    a.dst = av;
    a.aux = xv;
    a.sigma = sig.as<double>();
    a.partial = part.as<double>();
    // the fused epilogue's term with two vectors per thread where it is the LDS-staged one, else
    // one (the unstaged two-vector FP64 form is slower than two passes)
    const int nr2 = (nr == 2 && dslash_staged(g, 2, nt)) ? 2 : 1;
    launch_dslash(a, EPI_AXMY_DOT, nr2, nt, c.stream, tuned_cap(slot, KQ_MODE_FP64, nr2, nt));
    const uint32_t nb = blas_nblk(g.Vh);
    resid2_partial(b, ax.as<double>(), nl, list.as<int>(), g.Vh, part.as<double>(), nb, c.stream);
    const std::vector<double> all = reduce_host(part.as<double>(), nb, nk, c.stream);
    std::vector<double> out(nk, 0.0);
    for (int k : which)
        out[k] = all[k];
    KQ_HIP_CHECK(hipStreamSynchronize(c.stream));
    return out;
}

static std::vector<double> true_res2(int slot, int parity, const double *b, const double *x, int nk,
                                     const std::vector<double> &sigma_abs, std::vector<int> which = {})
{
    std::vector<double> out;
    with_fp64_links(slot, [&](LinkD fat, auto lng) {
        out = true_res2_t(slot, fat, lng, parity, b, x, nk, sigma_abs, which);
    });
    return out;
}

static bool debug_on()
{
    static const bool on = getenv("KQ_DEBUG") != nullptr;
    return on;
}

// autotuned right-hand-side blocking for 0..min(max_sys, 16) live systems
static std::vector<int> nr_table(int slot, int mode, int max_sys)
{
    const int n = std::min(max_sys, 16);
    std::vector<int> t(n + 1, 1);
    for (int k = 2; k <= n; ++k)
        t[k] = tuned_launch(slot, mode, k).first;
    return t;
}

static std::vector<double> abs_shifts(const KuQCDInvertParam &p, int r)
{
    // Q2: system k >= 1 is K + shifts[0] + shifts[k] (inverter_cgm.cpp:52)
    const double *s = p.shifts + (size_t)r * p.num_shift;
    std::vector<double> a(p.num_shift);
    for (int k = 0; k < p.num_shift; ++k)
        a[k] = (k == 0) ? s[0] : s[0] + s[k];
    return a;
}

// ---------------------------------------------------------------- mixed precision
struct Problem {
    int slot, parity, nrhs, ns;
    Geom g;
    const KuQCDInvertParam *p;
    const KqSolveOptions *o;
    DevBuf *b;  // FP64 [nrhs][3][Vh]
    std::vector<double> b2, tol, eps;
    double delta = 0.02;
};

// KQ_REFINE_L16=0 keeps the refinement of a 16-bit-link solve on its 16-bit fat links
static bool refine_l16_on()
{
    static const bool on = [] {
        const char *e = getenv("KQ_REFINE_L16");
        return !(e != nullptr && e[0] == '0');
    }();
    return on;
}

// Batched single-shift refinement of the listed (rhs, shift) pairs from their
// current FP64 solutions in x_out, to 0.5 * eps relative residual.
template <class LF, class LL>
static void refine_pairs(Problem &P, LF fat, LL lng, DevBuf &x_out, const std::vector<std::pair<int, int>> &need,
                         RunnerTuning tune, KqSolveReport *rep, int mode)
{
    Context &c = ctx();
    const size_t n3 = (size_t)3 * P.g.Vh;
    const int nsys = (int)need.size();
    if (P.o->nr > 0)
        tune.nr = P.o->nr;
    else
        tune.nr_by_live = nr_table(P.slot, mode, nsys);
    CgmRunner<double, float, LF, LL> ref(P.g, P.parity, fat, lng, nsys, nsys, c.stream, tune);
    with_fp64_links(P.slot, [&](LinkD f, auto l) { ref.enable_reliable(f, l, P.delta); });
    for (int i = 0; i < nsys; ++i) {
        const int r = need[i].first, k = need[i].second;
        const double sh = abs_shifts(*P.p, r)[k];
        const double tol = 0.25 * P.eps[r] * P.eps[r] * P.b2[r];
        ref.add_system(1, &sh, &tol, P.p->max_iter, P.b->as<double>() + (size_t)r * 2 * n3,
                       x_out.as<double>() + ((size_t)r * P.ns + k) * 2 * n3);
    }
    ref.run();
    for (int i = 0; i < nsys; ++i) {
        const int r = need[i].first, k = need[i].second;
        KQ_HIP_CHECK(hipMemcpyAsync(x_out.as<double>() + ((size_t)r * P.ns + k) * 2 * n3, ref.x_hi(i, 0),
                                    sizeof(double2) * n3, hipMemcpyDeviceToDevice, c.stream));
        rep[r].refine_iter[k] += ref.state(i).iter;
    }
    KQ_HIP_CHECK(hipStreamSynchronize(c.stream));
}

// Mixed precision: one multishift system per source in the sloppy format with
// FP64 reliable updates of the base shift. Each shift k >= 1 is forked out of
// the multishift as soon as its iterated relative residual reaches fork_tol
// and continues as an independent single-shift reliable CG from its FP64
// solution, inside the same batched iteration (the links are read once for all
// live systems). Every shift is then verified in FP64; any that misses the
// target is refined again.
template <class LF, class LL>
static void run_mixed(Problem &P, LF fat, LL lng, DevBuf &x_out, KqSolveReport *rep, std::vector<double> &res2)
{
    Context &c = ctx();
    const size_t n3 = (size_t)3 * P.g.Vh;
    const int ns = P.ns, nrhs = P.nrhs;
    RunnerTuning tune;
    tune.nt = P.o->nt > 0 ? P.o->nt : tuned_launch(P.slot, P.o->mode, nrhs).second;
    tune.chunk = P.o->chunk > 0 ? P.o->chunk : 8;
    // relative fork threshold of the iterated shifted residual. The true shifted residuals of
    // the sloppy multishift stop near 1e-6 (FP32) while the iterated ones keep falling; a shift
    // that stays until its iterated residual is 1e-7 costs only its x/p updates, and its fork
    // then starts later, beside fewer other forks (measured best at 32^4 for both formats)
    const double fork = P.o->fork_tol < 0 ? 0.0 : P.o->fork_tol > 0 ? P.o->fork_tol : 1e-7;
    const int max_sys = nrhs * ns;
    // right-hand-side blocking of the hopping term by number of live systems
    if (P.o->nr > 0)
        tune.nr = P.o->nr;
    else
        tune.nr_by_live = nr_table(P.slot, P.o->mode, max_sys);
    set_caps(tune, P.slot, P.o->mode, true);
    x_out.alloc(sizeof(double2) * nrhs * ns * n3);

    auto t0 = clk::now();
    CgmRunner<double, float, LF, LL> run(P.g, P.parity, fat, lng, max_sys, (size_t)nrhs * ns + (size_t)nrhs * (ns - 1),
                                         c.stream, tune);
    with_fp64_links(P.slot, [&](LinkD f, auto l) { run.enable_reliable(f, l, P.delta); });
    // the multishift systems fold their solution updates per segment instead of updating every
    // shift's x and p each iteration (KQ_DEFERRED=0 switches this off)
    run.enable_deferred(deferred_slots<float>(nrhs, P.g.Vh));
    std::vector<int> sys_of(nrhs);
    std::vector<int> rhs_of(max_sys, -1);
    std::vector<int> forked((size_t)nrhs * ns, -1);
    const bool learned = P.o->learned != 0;
    const double ltol = P.o->learned_tol > 0 ? P.o->learned_tol : 1e-6;
    for (int r = 0; r < nrhs; ++r) {
        std::vector<double> tolk(ns, P.tol[r]);
        if (fork > 0)
            for (int k = 1; k < ns; ++k)
                tolk[k] = std::max(P.tol[r], fork * fork * P.b2[r]);
        // shift 0 leaves for the flexible-CG refinement at learned_tol
        if (learned)
            tolk[0] = std::max(P.tol[r], ltol * ltol * P.b2[r]);
        sys_of[r] = run.add_system(ns, P.p->shifts + (size_t)r * ns, tolk.data(), P.p->max_iter,
                                   P.b->as<double>() + (size_t)r * 2 * n3, nullptr);
        rhs_of[sys_of[r]] = r;
    }
    auto on_drop = [&](int sys, int k) {
        const int r = rhs_of[sys];
        if (r < 0 || fork <= 0)
            return;
        const double sh = abs_shifts(*P.p, r)[k];
        const double tol = 0.25 * P.eps[r] * P.eps[r] * P.b2[r];
        forked[(size_t)r * ns + k] = run.add_system(1, &sh, &tol, P.p->max_iter, run.b_hi(sys), run.x_hi(sys, k));
    };
    const auto t_setup = clk::now();
    run.run(on_drop);
    const auto t_ran = clk::now();
    for (int r = 0; r < nrhs; ++r) {
        rep[r].iter = run.state(sys_of[r]).iter;
        rep[r].reliable_updates = run.reliable_updates();
        for (int k = 0; k < ns; ++k) {
            const int f = forked[(size_t)r * ns + k];
            const double *src = f >= 0 ? run.x_hi(f, 0) : run.x_hi(sys_of[r], k);
            if (f >= 0)
                rep[r].refine_iter[k] = run.state(f).iter;
            KQ_HIP_CHECK(hipMemcpyAsync(x_out.as<double>() + ((size_t)r * ns + k) * 2 * n3, src,
                                        sizeof(double2) * n3, hipMemcpyDeviceToDevice, c.stream));
        }
    }
    KQ_HIP_CHECK(hipStreamSynchronize(c.stream));
    auto t1 = clk::now();
    for (int r = 0; r < nrhs; ++r)
        rep[r].seconds_multishift = secs(t0, t1);

    // learned-preconditioner flexible CG for shift 0, from the multishift iterate
    double t_learned = 0.0;
    if (learned) {
        KqRefineOptions ro;
        kq_default_refine_options(&ro);
        ro.mode = P.o->mode;
        ro.learned = 1;
        ro.max_iter = P.p->max_iter > 0 ? P.p->max_iter : 1;
        ro.delta = P.delta;
        ro.verbose = P.o->verbose;
        for (int r = 0; r < nrhs; ++r) {
            ro.rel_tol = 0.9 * P.eps[r];
            const RefineResult rr = fcg_refine_device(P.slot, P.parity, abs_shifts(*P.p, r)[0],
                                                      P.b->as<double>() + (size_t)r * 2 * n3,
                                                      x_out.as<double>() + (size_t)r * ns * 2 * n3, ro);
            rep[r].refine_iter[0] = rr.iter;
            rep[r].learned_used = rr.fallback ? 0 : 1;
            rep[r].learned_fallback = rr.fallback ? 1 : 0;
            t_learned += rr.seconds;
        }
    }

    // per-shift FP64 verification; refine what misses the target
    res2.assign((size_t)nrhs * ns, 0.0);
    double t_verify = 0.0;
    auto verify = [&](const std::vector<std::pair<int, int>> &pairs) {
        auto tv0 = clk::now();
        for (int r = 0; r < nrhs; ++r) {
            std::vector<int> ks;
            for (const auto &pr : pairs)
                if (pr.first == r)
                    ks.push_back(pr.second);
            if (!pairs.empty() && ks.empty())
                continue;
            const auto v = true_res2(P.slot, P.parity, P.b->as<double>() + (size_t)r * 2 * n3,
                                     x_out.as<double>() + (size_t)r * ns * 2 * n3, ns, abs_shifts(*P.p, r), ks);
            for (int k = 0; k < ns; ++k)
                if (pairs.empty() || std::find(ks.begin(), ks.end(), k) != ks.end())
                    res2[(size_t)r * ns + k] = v[k];
        }
        t_verify += secs(tv0, clk::now());
    };
    verify({});
    for (int r = 0; r < nrhs; ++r)
        for (int k = 0; k < ns; ++k)
            rep[r].initial_rel_res[k] = std::sqrt(res2[(size_t)r * ns + k] / P.b2[r]);
    if (debug_on()) {
        fprintf(stderr, "[kq] mixed solve: setup %.2f ms, iterations %.2f ms, copy-out %.2f ms, verification %.2f ms\n",
                1e3 * secs(t0, t_setup), 1e3 * secs(t_setup, t_ran), 1e3 * secs(t_ran, t1), 1e3 * t_verify);
        for (int k = 0; k < ns; ++k)
            fprintf(stderr, "[kq] shift %d true rel res after the main phase %.3e (fork iters %d)\n", k,
                    rep[0].initial_rel_res[k], rep[0].refine_iter[k]);
    }
    auto t2 = clk::now();
    if (P.o->refine) {
        for (int round = 0; round < 3; ++round) {
            std::vector<std::pair<int, int>> need;
            for (int r = 0; r < nrhs; ++r)
                for (int k = 0; k < ns; ++k)
                    if (!(std::sqrt(res2[(size_t)r * ns + k] / P.b2[r]) <= P.eps[r]))
                        need.emplace_back(r, k);
            if (need.empty())
                break;
            if (P.o->mode == KQ_MODE_MIXED_FP16 && refine_l16_on()) {
                // 16-bit fat links misrepresent the lowest modes of the lightest systems (why
                // those stall): what the 16-bit solve left above the target is refined with FP32
                // fat links and the same 16-bit long links
                with_mixed_links(P.slot, KQ_MODE_MIXED_FP32_L16, [&](auto f, auto l) {
                    RunnerTuning t2;
                    t2.nt = P.o->nt > 0 ? P.o->nt : tuned_launch(P.slot, KQ_MODE_MIXED_FP32_L16, nrhs).second;
                    t2.chunk = tune.chunk;
                    set_caps(t2, P.slot, KQ_MODE_MIXED_FP32_L16, true);
                    refine_pairs<decltype(f), decltype(l)>(P, f, l, x_out, need, t2, rep, KQ_MODE_MIXED_FP32_L16);
                });
            } else {
                refine_pairs<LF, LL>(P, fat, lng, x_out, need, tune, rep, P.o->mode);
            }
            for (int r = 0; r < nrhs; ++r)
                rep[r].refine_rounds = round + 1;
            verify(need);
        }
    }
    auto t3 = clk::now();
    for (int r = 0; r < nrhs; ++r) {
        rep[r].seconds_refine = secs(t2, t3) - (t_verify - 0.0) + t_learned;
        rep[r].seconds_verify = t_verify;
    }
}

static void run_fp64(Problem &P, DevBuf &x_out, KqSolveReport *rep, std::vector<double> &res2)
{
    Context &c = ctx();
    const size_t n3 = (size_t)3 * P.g.Vh;
    RunnerTuning tune;
    tune.nt = P.o->nt > 0 ? P.o->nt : kReproNT;
    if (P.o->nr > 0)
        tune.nr = P.o->nr;
    else
        tune.nr_by_live = nr_table(P.slot, KQ_MODE_FP64, P.nrhs);
    set_caps(tune, P.slot, KQ_MODE_FP64, false);
    tune.chunk = P.o->chunk > 0 ? P.o->chunk : 16;
    auto t0 = clk::now();
    std::vector<int> iters(P.nrhs);
    x_out.alloc(sizeof(double2) * P.nrhs * P.ns * n3);
    with_fp64_links(P.slot, [&](auto fat, auto lng) {
        CgmRunner<double, double, decltype(fat), decltype(lng)> run(P.g, P.parity, fat, lng, P.nrhs,
                                                                     (size_t)P.nrhs * P.ns, c.stream, tune);
        run.enable_deferred(deferred_slots<double>(P.nrhs, P.g.Vh));
        for (int r = 0; r < P.nrhs; ++r) {
            std::vector<double> tolk(P.ns, P.tol[r]);
            run.add_system(P.ns, P.p->shifts + (size_t)r * P.ns, tolk.data(), P.p->max_iter,
                           P.b->as<double>() + (size_t)r * 2 * n3, nullptr);
        }
        run.run();
        for (int r = 0; r < P.nrhs; ++r) {
            iters[r] = run.state(r).iter;
            for (int k = 0; k < P.ns; ++k)
                KQ_HIP_CHECK(hipMemcpyAsync(x_out.as<double>() + ((size_t)r * P.ns + k) * 2 * n3, run.x_lo(r, k),
                                            sizeof(double2) * n3, hipMemcpyDeviceToDevice, c.stream));
        }
        KQ_HIP_CHECK(hipStreamSynchronize(c.stream));
    });
    auto t1 = clk::now();
    for (int r = 0; r < P.nrhs; ++r) {
        rep[r].iter = iters[r];
        rep[r].seconds_multishift = secs(t0, t1);
    }
    res2.assign((size_t)P.nrhs * P.ns, 0.0);
    auto tv0 = clk::now();
    if (P.o->refine)
        for (int r = 0; r < P.nrhs; ++r) {
            const auto v = true_res2(P.slot, P.parity, P.b->as<double>() + (size_t)r * 2 * n3,
                                     x_out.as<double>() + (size_t)r * P.ns * 2 * n3, P.ns, abs_shifts(*P.p, r));
            for (int k = 0; k < P.ns; ++k)
                res2[(size_t)r * P.ns + k] = v[k];
        }
    for (int r = 0; r < P.nrhs; ++r) {
        rep[r].seconds_verify = secs(tv0, clk::now());
        for (int k = 0; k < P.ns; ++k)
            rep[r].initial_rel_res[k] = std::sqrt(res2[(size_t)r * P.ns + k] / P.b2[r]);
    }
}

KuQCDStatus solve_extended(int slot, int nrhs, void **x, void **b, KuQCDInvertParam &p, const KqSolveOptions &o,
                           KqSolveReport *rep)
{
    auto t0 = clk::now();
    Context &c = ctx();
    std::vector<KqSolveReport> local(rep ? 0 : nrhs);
    if (!rep)
        rep = local.data();
    for (int r = 0; r < nrhs; ++r)
        std::memset(&rep[r], 0, sizeof(KqSolveReport));
    if (o.mode == KQ_MODE_REFERENCE && c.fat[slot].prec != KUQCD_DOUBLE) {
        if (nrhs != 1) {
            set_error("batched solves need a double-precision gauge field");
            return KUQCD_ERR_UNSUPPORTED;
        }
        KuQCDStatus st = solve_reference(slot, x, b[0], p);
        if (rep) {
            rep[0].iter = p.iter;
            rep[0].seconds = secs(t0, clk::now());
        }
        return st;
    }
    if (c.fat[slot].prec != KUQCD_DOUBLE) {
        set_error("the FP64 and mixed-precision modes need a double-precision gauge field");
        return KUQCD_ERR_UNSUPPORTED;
    }
    if (o.learned && o.mode != KQ_MODE_FP64 && o.mode != KQ_MODE_REFERENCE && !learned_weights_available()) {
        set_error("no preconditioner weights: call kq_ptnet_set_weights or kq_learned_train first");
        return KUQCD_ERR_UNSUPPORTED;
    }
    Problem P;
    P.slot = slot;
    P.parity = (p.matpc_type == KUQCD_EVEN_EVEN) ? 0 : 1;
    P.nrhs = nrhs;
    P.ns = p.num_shift;
    P.g = slot_geom(slot);
    P.p = &p;
    P.o = &o;
    // default interval: every 1.7 decades with FP32 fat links (3 % faster than every decade at
    // 32^4 and 48^3x96, same iteration counts); every decade with 16-bit fat links, whose
    // rounding lets the iterated residual drift faster (48^3x96: 1927-2123 iterations at 0.1,
    // up to 3842 at 0.02)
    P.delta = o.reliable_delta > 0 ? o.reliable_delta : (o.mode == KQ_MODE_MIXED_FP16 ? 0.1 : 0.02);
    const size_t n3 = (size_t)3 * P.g.Vh;
    DevBuf bdev(sizeof(double2) * nrhs * n3), stage;
    for (int r = 0; r < nrhs; ++r) {
        Vec<double> bv{bdev.as<double2>(), P.g.Vh, n3};
        upload_vec<double, double>(b[r], bv, r, stage, c.stream);
    }
    P.b = &bdev;
    {
        DevBuf part(sizeof(double) * nrhs * blas_nblk(P.g.Vh));
        Vec<double> bv{bdev.as<double2>(), P.g.Vh, n3};
        vec_dot_partial<double>(bv, bv, nrhs, part.as<double>(), blas_nblk(P.g.Vh), c.stream);
        P.b2 = reduce_host(part.as<double>(), blas_nblk(P.g.Vh), nrhs, c.stream);
    }
    P.tol.resize(nrhs);
    P.eps.resize(nrhs);
    for (int r = 0; r < nrhs; ++r) {
        if (o.rel_tol > 0) {
            P.tol[r] = o.rel_tol * o.rel_tol * P.b2[r];
            P.eps[r] = o.rel_tol;
        } else {
            P.tol[r] = p.tol;
            P.eps[r] = std::sqrt(p.tol / P.b2[r]);
        }
    }
    DevBuf xdev;
    std::vector<double> res2;
    const int mode = (o.mode == KQ_MODE_REFERENCE) ? KQ_MODE_FP64 : o.mode;
    if (mode == KQ_MODE_FP64) {
        run_fp64(P, xdev, rep, res2);
    } else if (is_mixed_mode(mode)) {
        with_mixed_links(slot, mode, [&](auto fat, auto lng) { run_mixed(P, fat, lng, xdev, rep, res2); });
    } else {
        set_error("unknown mode");
        return KUQCD_ERR_INVALID_PARAM;
    }
    // solutions to the host (double)
    for (int r = 0; r < nrhs; ++r)
        for (int k = 0; k < P.ns; ++k) {
            Vec<double> xv{xdev.as<double2>() + ((size_t)r * P.ns + k) * n3, P.g.Vh, n3};
            download_vec<double, double>(xv, 0, x[(size_t)r * P.ns + k], stage, c.stream);
        }
    const double total = secs(t0, clk::now());
    int maxit = 0;
    for (int r = 0; r < nrhs; ++r) {
        if (rep) {
            for (int k = 0; k < P.ns; ++k)
                rep[r].true_rel_res[k] = std::sqrt(res2[(size_t)r * P.ns + k] / P.b2[r]);
            rep[r].seconds = total;
            maxit = std::max(maxit, rep[r].iter);
        }
    }
    p.iter = maxit;
    if (o.verbose > 0)
        print_iters(maxit, p.max_iter);
    return KUQCD_SUCCESS;
}

}  // namespace kq
