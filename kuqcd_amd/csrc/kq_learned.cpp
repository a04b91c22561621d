// Learned preconditioner: the training set of solver-generated (residual,
// error) pairs, Adam training with the HIP forward/backward kernels, and the
// flexible-CG refinement (FP32 inner iteration, FP64 reliable updates, plain
// CG fallback) that uses the trained network.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <vector>

#include "kq_fcg.h"
#include "kq_gen.h"
#include "kq_learned.h"
#include "kq_ptnet.h"
#include "kq_solve.h"

namespace kq {

namespace {

using clk = std::chrono::steady_clock;
double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

// ---------------------------------------------------------------- state
struct TrainGauge {
    Geom g;
    DevBuf fat, lng;  // LinkF copies
    PtNorms nu;
};

struct TrainSample {
    int gauge = 0, parity = 0;
    double sigma = 0.0;
    double level = 0.0;  // relative residual of the iterate the pair was taken at
    DevBuf r, e, ae;     // FP32 SoA, |r| = 1; ae = A e with the training operator
};

struct Learned {
    bool have = false;
    PtWeights w{};
    std::vector<std::unique_ptr<TrainGauge>> gauges;
    std::vector<std::unique_ptr<TrainSample>> samples;
};

Learned &learned()
{
    static Learned L;
    return L;
}

PtWeights from_abi(const KqPtWeights &a)
{
    static_assert(sizeof(KqPtWeights) == sizeof(PtWeights), "weight layouts differ");
    PtWeights w;
    std::memcpy(&w, &a, sizeof w);
    return w;
}

KqPtWeights to_abi(const PtWeights &w)
{
    KqPtWeights a;
    std::memcpy(&a, &w, sizeof a);
    return a;
}

// counter-based normal deviates (SplitMix64 + Box-Muller)
struct Rng {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal()
    {
        const double u = uniform(), v = uniform();
        return std::sqrt(-2.0 * std::log(u)) * std::cos(6.283185307179586 * v);
    }
};

int slot_or_throw(bool need_double = true)
{
    const int slot = active_slot();
    if (slot < 0)
        throw std::runtime_error("no init gauge_smeared and gauge_naik");
    if (need_double && ctx().fat[slot].prec != KUQCD_DOUBLE)
        throw std::runtime_error("the learned preconditioner needs a double-precision gauge field");
    return slot;
}

PtNorms slot_norms(int slot)
{
    Context &c = ctx();
    const Geom g = slot_geom(slot);
    PtNorms nu;
    const GaugeSlot *s[2] = {&c.fat[slot], &c.lng[slot]};
    for (int h = 0; h < 2; ++h) {
        const double ss = vd_sumsq(s[h]->soa.as<double>(), s[h]->soa.bytes() / sizeof(double), c.stream);
        const double rows = 24.0 * g.Vh;  // 8 Vh links of 3 rows
        nu.nu[h] = ss > 0.0 ? std::sqrt(ss / rows) : 1.0;
    }
    return nu;
}

template <class Fn> void with_format(int slot, int mode, Fn &&fn) { (void)with_mixed_links(slot, mode, fn); }

// ---------------------------------------------------------------- A = sigma - D D
template <class LF, class LL, typename R> struct OpA {
    using V2 = c2_t<R>;
    Geom g;
    int q;
    LF fat;
    LL lng;
    int nt;
    int cap = 0;  // occupancy cap of the hopping term (waves per SIMD, 0 = none)
    uint32_t nblk() const { return dslash_nblk(g.Vh, nt); }
    // dst = sigma src - D D src (parity q), tmp on parity 1-q; part[nblk] = partials of Re<src, dst>
    void apply(const V2 *src, V2 *tmp, V2 *dst, const double *sigma, double *part, const int *run,
               hipStream_t st) const
    {
        DslashArgs<LF, LL, R> a{};
        a.g = g;
        a.fat = fat;
        a.lng = lng;
        a.nrhs = 1;
        a.run = run;
        a.dst_parity = 1 - q;
        a.src = Vec<R>{const_cast<V2 *>(src), g.Vh, 0};
        a.dst = Vec<R>{tmp, g.Vh, 0};
        launch_dslash(a, EPI_STORE, 1, nt, st, cap);
        a.dst_parity = q;
        a.src = Vec<R>{tmp, g.Vh, 0};
        a.aux = Vec<R>{const_cast<V2 *>(src), g.Vh, 0};
        a.dst = Vec<R>{dst, g.Vh, 0};
        a.sigma = sigma;
        a.partial = part;
        launch_dslash(a, EPI_AXMY_DOT, 1, nt, st, cap);
    }
};

// out = (sigma - D D) x in FP32 with the given links (allocates its own scratch)
template <class LF, class LL>
void apply_a_f32(const Geom &g, int q, LF fat, LL lng, double sigma, const float2 *x, float2 *out, int nt,
                 hipStream_t st)
{
    const size_t n3 = (size_t)3 * g.Vh;
    DevBuf t(sizeof(float2) * n3), sig(sizeof(double)), part(sizeof(double) * dslash_nblk(g.Vh, nt));
    KQ_HIP_CHECK(hipMemcpyAsync(sig.as<void>(), &sigma, sizeof(double), hipMemcpyHostToDevice, st));
    const OpA<LF, LL, float> A{g, q, fat, lng, nt};
    A.apply(x, t.as<float2>(), out, sig.as<double>(), part.as<double>(), nullptr, st);
    KQ_HIP_CHECK(hipStreamSynchronize(st));
}

// ---------------------------------------------------------------- flexible CG
struct FcgCfg {
    double sigma = 0.0;
    double tol2 = 0.0;  // stop when the FP64 |r|^2 is below it
    double delta = 0.1;
    int max_iter = 100000;
    int chunk = 8;
    int stall = 300;
    bool learned = false;
    PtWeightsBF wf{};
    int nt_a = 256, nt_d = 256;
    int cap_a = 0, cap_d = 0;
    int verbose = 0;
};

struct FcgOut {
    int iter = 0, iter_learned = 0, reliable = 0;
    bool fallback = false, converged = false;
    double rr0 = 0.0, rr = 0.0;
};

using FcgHook = std::function<void(const double2 *x64, const double2 *r64, double rr64)>;

template <class AF, class AL, class PF, class PL>
FcgOut fcg_solve(const Geom &g, int q, AF af, AL al, PF pf, PL pl, LinkD df, LinkD dl, const double2 *b64,
                 double2 *x64, const FcgCfg &cfg, const FcgHook &hook, hipStream_t st)
{
    const size_t n3 = (size_t)3 * g.Vh;
    const uint32_t nbv = fcg_nblk(n3), nbl = pt_layer_nblk(g.Vh);
    const OpA<AF, AL, float> A{g, q, af, al, cfg.nt_a, cfg.cap_a};
    const OpA<LinkD, LinkD, double> D64{g, q, df, dl, cfg.nt_d, cfg.cap_d};
    const size_t fb = sizeof(float2) * n3, db = sizeof(double2) * n3;
    DevBuf vr(fb), vz(cfg.learned ? fb : 0), vp(fb), vq(fb), vt(fb), vx(fb);
    // hidden channels between the two network layers as bfloat16 (48 B per site)
    DevBuf vh16(cfg.learned ? sizeof(kq_v4u) * n3 : 0);
    DevBuf ax64(db), t64(db), r64(hook ? db : 0), xbest(cfg.learned ? db : 0);
    DevBuf part_qz(sizeof(double) * 2 * std::max(nbv, nbl)), part_pr(sizeof(double) * 2 * nbv),
        part_pq(sizeof(double) * std::max(A.nblk(), D64.nblk())), part_rr(sizeof(double) * nbv);
    DevBuf sdev(sizeof(FcgState));
    Pinned<FcgState> sh(1);
    FcgState &H = *sh.get();
    FcgState *S = sdev.as<FcgState>();
    std::memset(&H, 0, sizeof H);
    H.sigma = cfg.sigma;
    H.first = 1;
    H.stop = 1;
    H.max_iter = cfg.max_iter;
    H.rr_stop = cfg.tol2;
    auto upload = [&] {
        copy_words(sh.dev(), S, sizeof H, st);
    };
    upload();
    KQ_HIP_CHECK(hipMemsetAsync(vx.as<void>(), 0, fb, st));
    KQ_HIP_CHECK(hipMemsetAsync(vp.as<void>(), 0, fb, st));
    KQ_HIP_CHECK(hipMemsetAsync(vq.as<void>(), 0, fb, st));

    // x64 += x32; r = b - A x64 in FP64; returns |r|^2
    auto reliable = [&]() -> double {
        fcg_fold(x64, vx.as<float2>(), n3, st);
        D64.apply(x64, t64.as<double2>(), ax64.as<double2>(), &S->sigma, part_pq.as<double>(), nullptr, st);
        fcg_residual(b64, ax64.as<double2>(), r64 ? r64.as<double2>() : nullptr, vr.as<float2>(), n3,
                     part_rr.as<double>(), st);
        return reduce_host(part_rr.as<double>(), nbv, 1, st)[0];
    };

    auto enqueue = [&](bool net) {
        if (net) {
            pt_forward<PF, PL>(g, q, pf, pl, cfg.wf, vr.as<float2>(), nullptr, vz.as<float2>(), st,
                               vq.as<float2>(), part_qz.as<double>(), &S->stop, vh16.as<kq_v4u>());
            fcg_beta(S, part_qz.as<double>(), nbl, st);
            fcg_p_update(S, vz.as<float2>(), vp.as<float2>(), vr.as<float2>(), n3, part_pr.as<double>(), st);
        } else {
            fcg_dot(S, vq.as<float2>(), vr.as<float2>(), n3, part_qz.as<double>(), st);
            fcg_beta(S, part_qz.as<double>(), nbv, st);
            fcg_p_update(S, vr.as<float2>(), vp.as<float2>(), vr.as<float2>(), n3, part_pr.as<double>(), st);
        }
        A.apply(vp.as<float2>(), vt.as<float2>(), vq.as<float2>(), &S->sigma, part_pq.as<double>(), &S->run, st);
        fcg_alpha(S, part_pq.as<double>(), A.nblk(), part_pr.as<double>(), nbv, st);
        fcg_xr_update(S, vx.as<float2>(), vr.as<float2>(), vp.as<float2>(), vq.as<float2>(), n3,
                      part_rr.as<double>(), st);
        fcg_check(S, part_rr.as<double>(), nbv, st);
    };
    hipGraphExec_t graphs[2] = {nullptr, nullptr};
    struct GraphGuard {
        hipGraphExec_t *g;
        ~GraphGuard()
        {
            for (int i = 0; i < 2; ++i)
                if (g[i])
                    (void)hipGraphExecDestroy(g[i]);
        }
    } guard{graphs};
    auto graph_for = [&](bool net) {
        hipGraphExec_t &gx = graphs[net ? 1 : 0];
        if (!gx) {
            hipGraph_t graph;
            KQ_HIP_CHECK(hipStreamSynchronize(st));
            KQ_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            for (int i = 0; i < cfg.chunk; ++i)
                enqueue(net);
            KQ_HIP_CHECK(hipStreamEndCapture(st, &graph));
            KQ_HIP_CHECK(hipGraphInstantiate(&gx, graph, nullptr, nullptr, 0));
            KQ_HIP_CHECK(hipGraphDestroy(graph));
        }
        return gx;
    };

    FcgOut out;
    double rr = reliable();
    out.rr0 = rr;
    if (hook)
        hook(x64, r64.as<double2>(), rr);
    bool net = cfg.learned;
    double best = rr;
    if (net)
        KQ_HIP_CHECK(hipMemcpyAsync(xbest.as<void>(), x64, db, hipMemcpyDeviceToDevice, st));
    for (;;) {
        if (std::isnan(rr))
            break;
        if (rr <= cfg.tol2) {
            out.converged = true;
            break;
        }
        if (H.iter >= cfg.max_iter)
            break;
        H.rr = rr;
        H.rr_pause = cfg.delta * cfg.delta * rr;
        H.stop = 0;
        H.run = 1;
        H.reason = 0;
        H.max_iter = net ? std::min(cfg.max_iter, H.iter + cfg.stall) : cfg.max_iter;
        upload();
        const int it0 = H.iter;
        hipGraphExec_t gx = direct_launches() ? nullptr : graph_for(net);
        do {
            if (gx)
                KQ_HIP_CHECK(hipGraphLaunch(gx, st));
            else
                for (int i = 0; i < cfg.chunk; ++i)
                    enqueue(net);
            copy_words(S, sh.dev(), sizeof H, st);
            KQ_HIP_CHECK(hipStreamSynchronize(st));
        } while (!H.stop);
        if (net)
            out.iter_learned += H.iter - it0;
        const int reason = H.reason;
        rr = reliable();
        ++out.reliable;
        if (hook)
            hook(x64, r64.as<double2>(), rr);
        if (cfg.verbose > 1)
            log_info("[kq] fcg%s: iter %d, |r|^2 %.4e (inner %.4e), reason %d\n", net ? " (learned)" : "", H.iter, rr,
                     H.rr, reason);
        if (net) {
            const bool worse = !(rr < best);
            const bool stalled = reason == kFcgStopMaxIter && H.iter < cfg.max_iter;
            if (worse || stalled) {
                // the network did not reduce the FP64 residual: plain CG from the best iterate
                net = false;
                out.fallback = true;
                H.first = 1;
                if (worse) {
                    KQ_HIP_CHECK(hipMemcpyAsync(x64, xbest.as<void>(), db, hipMemcpyDeviceToDevice, st));
                    rr = reliable();
                }
                if (cfg.verbose > 0)
                    log_info("[kq] learned refinement rejected after %d iterations; plain CG\n", H.iter);
                continue;
            }
            KQ_HIP_CHECK(hipMemcpyAsync(xbest.as<void>(), x64, db, hipMemcpyDeviceToDevice, st));
        }
        best = std::min(best, rr);
    }
    KQ_HIP_CHECK(hipStreamSynchronize(st));
    out.iter = H.iter;
    out.rr = rr;
    return out;
}

// ---------------------------------------------------------------- loss and gradient
constexpr int kTrainNT = 128;  // hopping-term block size of the training kernels

struct Work {
    DevBuf h, z, az, t, d, gz, gh, sig, part;
    size_t n3 = 0;
    int nt = 0;
    void ensure(uint32_t Vh, int nt_)
    {
        const size_t n = (size_t)3 * Vh;
        if (n <= n3 && nt_ == nt)
            return;
        n3 = n;
        nt = nt_;
        const size_t fb = sizeof(float2) * n;
        h.alloc(4 * fb);
        z.alloc(fb);
        az.alloc(fb);
        t.alloc(fb);
        d.alloc(fb);
        gz.alloc(fb);
        gh.alloc(4 * fb);
        sig.alloc(sizeof(double));
        part.alloc(sizeof(double) * dslash_nblk(Vh, nt));
    }
};

// loss L = |A(z - e)|^2 / |A e|^2 with z = M r, for a pair (r, e) with A e = ae given:
// the error of the correction z in the A^2 norm (for a solver pair, r = A e and L is
// the residual one application of M leaves). grad (optional, 2*kPtParams doubles):
// dL/dconj(w) of the master weights, from dL/dconj(z) = A (A z - A e) / |A e|^2.
template <class LF, class LL>
double sample_loss(const Geom &g, int q, LF fat, LL lng, const PtNorms &nu, const PtWeights &w, double sigma,
                   const float2 *r, const float2 *ae, Work &W, int nt, double *grad, hipStream_t st)
{
    W.ensure(g.Vh, nt);
    const size_t n3 = (size_t)3 * g.Vh;
    pt_forward<LF, LL>(g, q, fat, lng, pt_pack(w, nu), r, W.h.as<float2>(), W.z.as<float2>(), st);
    KQ_HIP_CHECK(hipMemcpyAsync(W.sig.as<void>(), &sigma, sizeof(double), hipMemcpyHostToDevice, st));
    const OpA<LF, LL, float> A{g, q, fat, lng, nt};
    A.apply(W.z.as<float2>(), W.t.as<float2>(), W.az.as<float2>(), W.sig.as<double>(), W.part.as<double>(), nullptr,
            st);
    vf_lincomb(W.d.as<float2>(), W.az.as<float2>(), mk<float2>(1.f, 0.f), ae, mk<float2>(-1.f, 0.f), n3, st);
    const float2 *xs[2] = {W.d.as<float2>(), ae};
    const float2 *ys[2] = {W.d.as<float2>(), ae};
    double2 dd[2];
    vf_dots(xs, ys, 2, n3, dd, st);
    const double num = dd[0].x, den = dd[1].x;
    if (grad)
        std::fill(grad, grad + 2 * kPtParams, 0.0);
    if (!(den > 0.0))
        return 0.0;
    const double loss = num / den;
    if (!grad)
        return loss;
    A.apply(W.d.as<float2>(), W.t.as<float2>(), W.gz.as<float2>(), W.sig.as<double>(), W.part.as<double>(), nullptr,
            st);
    const float inv = (float)(1.0 / den);
    vf_lincomb(W.gz.as<float2>(), W.gz.as<float2>(), mk<float2>(inv, 0.f), W.gz.as<float2>(), mk<float2>(0.f, 0.f),
               n3, st);
    pt_backward_hidden<LF, LL>(g, q, fat, lng, pt_pack_backward(w, nu), W.gz.as<float2>(), W.gh.as<float2>(), st);
    pt_weight_grads<LF, LL>(g, q, fat, lng, r, W.h.as<float2>(), W.gz.as<float2>(), W.gh.as<float2>(), grad, st);
    // w_eff = w / nu_hop for the tap weights
    for (int l = 0; l < 2; ++l)
        for (int c = 0; c < kPtHidden; ++c)
            for (int t = 0; t < kPtTaps; ++t) {
                const int j = l * kPtHidden * kPtTaps + c * kPtTaps + t;
                const double sc = 1.0 / nu.nu[pt_tap_hop(t) == 3 ? 1 : 0];
                grad[2 * j] *= sc;
                grad[2 * j + 1] *= sc;
            }
    return loss;
}

double train_sample_loss(const TrainSample &s, const PtWeights &w, Work &W, double *grad, hipStream_t st)
{
    const TrainGauge &G = *learned().gauges[s.gauge];
    const LinkF fat = as_linkF(G.fat, G.g.Vh), lng = as_linkF(G.lng, G.g.Vh);
    return sample_loss<LinkF, LinkF>(G.g, s.parity, fat, lng, G.nu, w, s.sigma, s.r.as<float2>(), s.ae.as<float2>(),
                                     W, kTrainNT, grad, st);
}

double mean_loss(const PtWeights &w, Work &W, hipStream_t st)
{
    const Learned &L = learned();
    double s = 0.0;
    for (const auto &p : L.samples)
        s += train_sample_loss(*p, w, W, nullptr, st);
    return L.samples.empty() ? 0.0 : s / (double)L.samples.size();
}

void init_weights(PtWeights &w, uint64_t seed, double scale)
{
    Rng rng{seed * 0x2545f4914f6cdd1dull + 0x1234567ull};
    for (int c = 0; c < kPtHidden; ++c)
        for (int t = 0; t < kPtTaps; ++t) {
            w.w1[c][t] = mk<float2>((float)(scale * rng.normal()), (float)(scale * rng.normal()));
            w.w2[c][t] = mk<float2>((float)(scale * rng.normal()), (float)(scale * rng.normal()));
        }
    w.w0 = mk<float2>(1.f, 0.f);
}

// Largest eigenvalue of A = sigma - D D on the first training gauge (power iteration, FP32).
double lambda_max_train(int iters, hipStream_t st)
{
    const Learned &L = learned();
    const TrainGauge &G = *L.gauges[L.samples[0]->gauge];
    const int q = L.samples[0]->parity;
    const double sigma = L.samples[0]->sigma;
    const size_t n3 = (size_t)3 * G.g.Vh;
    std::vector<float2> h(n3);
    Rng rng{0x5eedull};
    for (auto &v : h)
        v = mk<float2>((float)rng.normal(), (float)rng.normal());
    DevBuf y(sizeof(float2) * n3), ay(sizeof(float2) * n3);
    KQ_HIP_CHECK(hipMemcpyAsync(y.as<void>(), h.data(), sizeof(float2) * n3, hipMemcpyHostToDevice, st));
    const LinkF fat = as_linkF(G.fat, G.g.Vh), lng = as_linkF(G.lng, G.g.Vh);
    double lam = 0.0;
    for (int i = 0; i < iters; ++i) {
        apply_a_f32(G.g, q, fat, lng, sigma, y.as<float2>(), ay.as<float2>(), kTrainNT, st);
        const float2 *xs[2] = {y.as<float2>(), ay.as<float2>()};
        const float2 *ys[2] = {y.as<float2>(), ay.as<float2>()};
        double2 d[2];
        vf_dots(xs, ys, 2, n3, d, st);
        lam = std::sqrt(d[1].x / d[0].x);
        const float inv = (float)(1.0 / std::sqrt(d[1].x));
        vf_lincomb(y.as<float2>(), ay.as<float2>(), mk<float2>(inv, 0.f), ay.as<float2>(), mk<float2>(0.f, 0.f), n3,
                   st);
    }
    return lam;
}

// Start from the degree-1 polynomial of A that a Chebyshev bound suggests,
// M = 8/beta - 0.9 * 8/beta^2 * A (beta = 1.02 lambda_max, positive on the spectrum),
// written in channel 0 (h_0 = D r, then D h_0 = D D r = sigma - A); the other
// channels get small random weights.
void init_chebyshev(PtWeights &w, uint64_t seed, double noise, hipStream_t st)
{
    const Learned &L = learned();
    const PtNorms &nu = L.gauges[L.samples[0]->gauge]->nu;
    const double sigma = L.samples[0]->sigma;
    const double beta = 1.02 * lambda_max_train(40, st);
    const double c = 0.9 * 8.0 / (beta * beta);
    init_weights(w, seed, noise);
    for (int t = 0; t < kPtTaps; ++t) {
        const double d = (t & 2) ? -1.0 : 1.0;  // sign of tap t in the hopping term
        const double n = nu.nu[pt_tap_hop(t) == 3 ? 1 : 0];
        w.w1[0][t] = mk<float2>((float)(d * n), 0.f);
        w.w2[0][t] = mk<float2>((float)(c * d * n), 0.f);
    }
    w.w0 = mk<float2>((float)(8.0 / beta - c * sigma), 0.f);
}

int tuned_nt(int slot, int mode)
{
    return tuned_launch(slot, mode, 1).second;
}

template <class Fn> KuQCDStatus guarded(Fn fn)
{
    try {
        return fn();
    } catch (const std::bad_alloc &) {
        set_error("out of memory");
        return KUQCD_ERR_NO_MEMORY;
    } catch (const std::invalid_argument &e) {
        set_error(e.what());
        return KUQCD_ERR_INVALID_PARAM;
    } catch (const std::exception &e) {
        set_error(e.what());
        return KUQCD_ERR_UNKNOWN_ERROR;
    }
}

}  // namespace

bool learned_weights_available() { return learned().have; }

void learned_release()
{
    Learned &L = learned();
    L.samples.clear();
    L.gauges.clear();
}

RefineResult fcg_refine_device(int slot, int parity, double sigma, const double *b, double *x,
                               const KqRefineOptions &o)
{
    Context &c = ctx();
    const Geom g = slot_geom(slot);
    const size_t n3 = (size_t)3 * g.Vh;
    const bool use_net = o.learned != 0;
    if (use_net && !learned().have)
        throw std::runtime_error("no preconditioner weights: call kq_ptnet_set_weights or kq_learned_train first");
    const int amode = o.mode == KQ_MODE_MIXED_FP16 ? KQ_MODE_MIXED_FP16 : KQ_MODE_MIXED_FP32;
    const int nmode = o.net_format == KQ_MODE_MIXED_FP32 ? KQ_MODE_MIXED_FP32 : KQ_MODE_MIXED_FP16;
    const double b2 = vd_sumsq(b, 2 * n3, c.stream);
    FcgCfg cfg;
    cfg.sigma = sigma;
    cfg.tol2 = o.rel_tol * o.rel_tol * b2;
    cfg.delta = o.delta > 0 ? o.delta : 0.1;
    cfg.max_iter = o.max_iter > 0 ? o.max_iter : 100000;
    cfg.chunk = o.chunk > 0 ? o.chunk : 8;
    cfg.stall = o.stall_iter > 0 ? o.stall_iter : 300;
    cfg.learned = use_net;
    cfg.verbose = o.verbose;
    cfg.nt_a = tuned_nt(slot, amode);
    cfg.nt_d = tuned_nt(slot, KQ_MODE_FP64);
    cfg.cap_a = tuned_cap(slot, amode, 1, cfg.nt_a);
    cfg.cap_d = tuned_cap(slot, KQ_MODE_FP64, 1, cfg.nt_d);
    if (use_net)
        cfg.wf = pt_pack(learned().w, slot_norms(slot));
    const LinkD df = link_d(c.fat[slot]), dl = link_d(c.lng[slot]);
    FcgOut out;
    const auto t0 = clk::now();
    with_format(slot, amode, [&](auto af, auto al) {
        with_format(slot, nmode, [&](auto pf, auto pl) {
            out = fcg_solve(g, parity, af, al, pf, pl, df, dl, reinterpret_cast<const double2 *>(b),
                            reinterpret_cast<double2 *>(x), cfg, nullptr, c.stream);
        });
    });
    RefineResult r;
    r.seconds = secs(t0, clk::now());
    r.iter = out.iter;
    r.iter_learned = out.iter_learned;
    r.reliable = out.reliable;
    r.fallback = out.fallback;
    r.converged = out.converged;
    r.rel0 = b2 > 0 ? std::sqrt(out.rr0 / b2) : 0.0;
    r.rel = b2 > 0 ? std::sqrt(out.rr / b2) : 0.0;
    return r;
}

}  // namespace kq

using namespace kq;

extern "C" {

void kq_ptnet_init_weights(KqPtWeights *w, uint64_t seed, double scale)
{
    if (!w)
        return;
    PtWeights p{};
    init_weights(p, seed, scale);
    *w = to_abi(p);
}

KuQCDStatus kq_ptnet_set_weights(const KqPtWeights *w)
{
    if (!w) {
        set_error("w is nullptr");
        return KUQCD_ERR_NULL_POINTER;
    }
    learned().w = from_abi(*w);
    learned().have = true;
    return KUQCD_SUCCESS;
}

KuQCDStatus kq_ptnet_get_weights(KqPtWeights *w)
{
    if (!w) {
        set_error("w is nullptr");
        return KUQCD_ERR_NULL_POINTER;
    }
    if (!learned().have) {
        set_error("no preconditioner weights set or trained");
        return KUQCD_ERR_UNSUPPORTED;
    }
    *w = to_abi(learned().w);
    return KUQCD_SUCCESS;
}

KuQCDStatus kq_ptnet_norms(double nu[2])
{
    if (!nu) {
        set_error("nu is nullptr");
        return KUQCD_ERR_NULL_POINTER;
    }
    return guarded([&] {
        const PtNorms n = slot_norms(slot_or_throw());
        nu[0] = n.nu[0];
        nu[1] = n.nu[1];
        return KUQCD_SUCCESS;
    });
}

static KuQCDStatus ptnet_forward(const KqPtWeights *w, int parity, int link_format, const void *r, void *z, void *h,
                                 bool hidden16)
{
    if (!w || !r || !z) {
        set_error("w, r and z must not be nullptr");
        return KUQCD_ERR_NULL_POINTER;
    }
    if (parity != 0 && parity != 1) {
        set_error("parity must be 0 or 1");
        return KUQCD_ERR_INVALID_PARAM;
    }
    return guarded([&] {
        Context &c = ctx();
        const int slot = slot_or_throw();
        const Geom g = slot_geom(slot);
        const size_t n3 = (size_t)3 * g.Vh;
        const PtWeightsBF wf = pt_pack(from_abi(*w), slot_norms(slot));
        DevBuf vr(sizeof(float2) * n3), vz(sizeof(float2) * n3), vh(sizeof(float2) * 4 * n3), stage;
        DevBuf vh16(hidden16 ? sizeof(kq_v4u) * n3 : 0);
        upload_vec<double, float>(r, Vec<float>{vr.as<float2>(), g.Vh, n3}, 0, stage, c.stream);
        with_format(slot, link_format, [&](auto fat, auto lng) {
            pt_forward(g, parity, fat, lng, wf, vr.as<float2>(), hidden16 ? nullptr : vh.as<float2>(),
                       vz.as<float2>(), c.stream, nullptr, nullptr, nullptr, hidden16 ? vh16.as<kq_v4u>() : nullptr);
        });
        if (hidden16 && h)
            pt_unpack_hidden(vh16.as<kq_v4u>(), vh.as<float2>(), g.Vh, c.stream);
        download_vec<float, double>(Vec<float>{vz.as<float2>(), g.Vh, n3}, 0, z, stage, c.stream);
        if (h)
            for (int ch = 0; ch < kPtHidden; ++ch)
                download_vec<float, double>(Vec<float>{vh.as<float2>(), g.Vh, n3}, ch,
                                            static_cast<double *>(h) + (size_t)ch * 2 * n3, stage, c.stream);
        return KUQCD_SUCCESS;
    });
}

KuQCDStatus kq_ptnet_forward(const KqPtWeights *w, int parity, int link_format, const void *r, void *z, void *h)
{
    return ptnet_forward(w, parity, link_format, r, z, h, false);
}

KuQCDStatus kq_ptnet_forward_h16(const KqPtWeights *w, int parity, int link_format, const void *r, void *z, void *h)
{
    return ptnet_forward(w, parity, link_format, r, z, h, true);
}

KuQCDStatus kq_ptnet_loss_grad(const KqPtWeights *w, int parity, double sigma, const void *r, const void *e,
                               double *loss, double *grad)
{
    if (!w || !r || !e || !loss) {
        set_error("w, r, e and loss must not be nullptr");
        return KUQCD_ERR_NULL_POINTER;
    }
    if (parity != 0 && parity != 1) {
        set_error("parity must be 0 or 1");
        return KUQCD_ERR_INVALID_PARAM;
    }
    return guarded([&] {
        Context &c = ctx();
        const int slot = slot_or_throw();
        const Geom g = slot_geom(slot);
        const size_t n3 = (size_t)3 * g.Vh;
        DevBuf vr(sizeof(float2) * n3), ve(sizeof(float2) * n3), vae(sizeof(float2) * n3), stage;
        upload_vec<double, float>(r, Vec<float>{vr.as<float2>(), g.Vh, n3}, 0, stage, c.stream);
        upload_vec<double, float>(e, Vec<float>{ve.as<float2>(), g.Vh, n3}, 0, stage, c.stream);
        DerivedFormats &d = derived(slot, true, false);
        const LinkF fat = as_linkF(d.fat_f, g.Vh), lng = as_linkF(d.lng_f, g.Vh);
        apply_a_f32(g, parity, fat, lng, sigma, ve.as<float2>(), vae.as<float2>(), kTrainNT, c.stream);
        Work W;
        *loss = sample_loss<LinkF, LinkF>(g, parity, fat, lng, slot_norms(slot), from_abi(*w), sigma, vr.as<float2>(),
                                          vae.as<float2>(), W, kTrainNT, grad, c.stream);
        return KUQCD_SUCCESS;
    });
}

KuQCDStatus kq_learned_collect(int parity, double sigma, int nsrc, uint64_t seed, double rel_tol, double min_level,
                               int *nsamples)
{
    if (parity != 0 && parity != 1) {
        set_error("parity must be 0 or 1");
        return KUQCD_ERR_INVALID_PARAM;
    }
    if (nsrc < 0 || !(rel_tol > 0.0) || !(rel_tol < 1.0)) {
        set_error("nsrc must be >= 0 and rel_tol in (0, 1)");
        return KUQCD_ERR_INVALID_PARAM;
    }
    return guarded([&] {
        Context &c = ctx();
        Learned &L = learned();
        const int slot = slot_or_throw();
        const Geom g = slot_geom(slot);
        const size_t n3 = (size_t)3 * g.Vh;
        DerivedFormats &d = derived(slot, true, false);
        auto tg = std::make_unique<TrainGauge>();
        tg->g = g;
        tg->fat.alloc(d.fat_f.bytes());
        tg->lng.alloc(d.lng_f.bytes());
        KQ_HIP_CHECK(hipMemcpyAsync(tg->fat.as<void>(), d.fat_f.as<void>(), d.fat_f.bytes(), hipMemcpyDeviceToDevice,
                                    c.stream));
        KQ_HIP_CHECK(hipMemcpyAsync(tg->lng.as<void>(), d.lng_f.as<void>(), d.lng_f.bytes(), hipMemcpyDeviceToDevice,
                                    c.stream));
        tg->nu = slot_norms(slot);
        const int gi = (int)L.gauges.size();
        L.gauges.push_back(std::move(tg));

        FcgCfg cfg;
        cfg.sigma = sigma;
        cfg.nt_a = tuned_nt(slot, KQ_MODE_MIXED_FP32);
        cfg.nt_d = tuned_nt(slot, KQ_MODE_FP64);
        const LinkF af = as_linkF(d.fat_f, g.Vh), al = as_linkF(d.lng_f, g.Vh);
        const LinkD df = link_d(c.fat[slot]), dl = link_d(c.lng[slot]);
        const unsigned *lat = c.fat[slot].lat;
        std::vector<double> bh((size_t)2 * n3);
        DevBuf b64(sizeof(double2) * n3), x64(sizeof(double2) * n3), stage;
        for (int j = 0; j < nsrc; ++j) {
            gen_source<double>(bh.data(), lat, parity, seed + (uint64_t)j);
            double b2 = 0.0;
            for (double v : bh)
                b2 += v * v;
            upload_vec<double, double>(bh.data(), Vec<double>{b64.as<double2>(), g.Vh, n3}, 0, stage, c.stream);
            KQ_HIP_CHECK(hipMemsetAsync(x64.as<void>(), 0, sizeof(double2) * n3, c.stream));
            std::vector<DevBuf> xs, rs;
            std::vector<double> lv;
            cfg.tol2 = rel_tol * rel_tol * b2;
            auto hook = [&](const double2 *x, const double2 *r, double rr) {
                const double level = std::sqrt(rr / b2);
                if (!(level >= min_level) || level > 1.0 + 1e-12)
                    return;
                xs.emplace_back(sizeof(double2) * n3);
                rs.emplace_back(sizeof(double2) * n3);
                KQ_HIP_CHECK(hipMemcpyAsync(xs.back().as<void>(), x, sizeof(double2) * n3, hipMemcpyDeviceToDevice,
                                            c.stream));
                KQ_HIP_CHECK(hipMemcpyAsync(rs.back().as<void>(), r, sizeof(double2) * n3, hipMemcpyDeviceToDevice,
                                            c.stream));
                lv.push_back(level);
            };
            const FcgOut out = fcg_solve(g, parity, af, al, af, al, df, dl, b64.as<double2>(), x64.as<double2>(), cfg,
                                         hook, c.stream);
            if (!out.converged)
                throw std::runtime_error("training solve did not converge");
            for (size_t k = 0; k < xs.size(); ++k) {
                auto s = std::make_unique<TrainSample>();
                s->gauge = gi;
                s->parity = parity;
                s->sigma = sigma;
                s->level = lv[k];
                s->r.alloc(sizeof(float2) * n3);
                s->e.alloc(sizeof(float2) * n3);
                const double scale = 1.0 / (lv[k] * std::sqrt(b2));
                fcg_to_f32(rs[k].as<double2>(), nullptr, scale, s->r.as<float2>(), n3, c.stream);
                fcg_to_f32(x64.as<double2>(), xs[k].as<double2>(), scale, s->e.as<float2>(), n3, c.stream);
                s->ae.alloc(sizeof(float2) * n3);
                apply_a_f32(g, parity, af, al, sigma, s->e.as<float2>(), s->ae.as<float2>(), kTrainNT, c.stream);
                L.samples.push_back(std::move(s));
            }
            KQ_HIP_CHECK(hipStreamSynchronize(c.stream));
            log_info("[kq] learned: source %d: %zu pairs, %d iterations\n", j, xs.size(), out.iter);
        }
        if (nsamples)
            *nsamples = (int)L.samples.size();
        return KUQCD_SUCCESS;
    });
}

void kq_learned_clear(void)
{
    try {
        Context &c = ctx();
        if (c.device_ready)
            (void)hipStreamSynchronize(c.stream);
        learned_release();
    } catch (...) {
    }
}

KuQCDStatus kq_learned_eval(const KqPtWeights *w, double *mean)
{
    if (!w || !mean) {
        set_error("w and mean_loss must not be nullptr");
        return KUQCD_ERR_NULL_POINTER;
    }
    return guarded([&] {
        if (learned().samples.empty())
            throw std::runtime_error("the training set is empty");
        Work W;
        *mean = mean_loss(from_abi(*w), W, ctx().stream);
        return KUQCD_SUCCESS;
    });
}

KuQCDStatus kq_learned_train(const KqTrainOptions *o, KqTrainReport *rep)
{
    if (!o) {
        set_error("options are nullptr");
        return KUQCD_ERR_NULL_POINTER;
    }
    if (o->steps < 0 || o->batch <= 0 || !(o->lr > 0.0)) {
        set_error("steps >= 0, batch > 0 and lr > 0 are required");
        return KUQCD_ERR_INVALID_PARAM;
    }
    return guarded([&] {
        Learned &L = learned();
        Context &c = ctx();
        if (L.samples.empty())
            throw std::runtime_error("the training set is empty: call kq_learned_collect first");
        const auto t0 = clk::now();
        PtWeights w = L.w;
        const double noise = o->init_scale > 0 ? o->init_scale : 0.01;
        if (o->init == 1)
            init_weights(w, o->seed, noise);
        else if (o->init == 2 || !L.have)
            init_chebyshev(w, o->seed, noise, c.stream);
        Work W;
        KqTrainReport R;
        std::memset(&R, 0, sizeof R);
        R.nsamples = (int)L.samples.size();
        R.ngauges = (int)L.gauges.size();
        R.loss_initial = mean_loss(w, W, c.stream);
        constexpr int P = 2 * kPtParams;
        float *prm = reinterpret_cast<float *>(&w);
        std::vector<double> m(P, 0.0), v(P, 0.0), gacc(P), gs(P);
        const double b1 = 0.9, b2 = 0.999, eps = 1e-12;
        const double lr_final = o->lr_final > 0 ? o->lr_final : 0.1 * o->lr;
        Rng rng{o->seed ^ 0x5bd1e995ull};
        std::vector<int> order(L.samples.size());
        size_t pos = order.size();
        int hist_n = 0;
        for (int step = 0; step < o->steps; ++step) {
            std::fill(gacc.begin(), gacc.end(), 0.0);
            double lsum = 0.0;
            for (int b = 0; b < o->batch; ++b) {
                if (pos >= order.size()) {  // new epoch: shuffled order
                    for (size_t i = 0; i < order.size(); ++i)
                        order[i] = (int)i;
                    for (size_t i = order.size(); i > 1; --i)
                        std::swap(order[i - 1], order[rng.next() % i]);
                    pos = 0;
                }
                lsum += train_sample_loss(*L.samples[order[pos++]], w, W, gs.data(), c.stream);
                for (int k = 0; k < P; ++k)
                    gacc[k] += gs[k];
            }
            const double lr = lr_final + 0.5 * (o->lr - lr_final) * (1.0 + std::cos(3.141592653589793 * step / std::max(1, o->steps)));
            const double bc1 = 1.0 - std::pow(b1, step + 1), bc2 = 1.0 - std::pow(b2, step + 1);
            for (int k = 0; k < P; ++k) {
                // real gradient of a real loss: 2 Re, 2 Im of dL/dconj(w)
                const double gk = 2.0 * gacc[k] / o->batch;
                m[k] = b1 * m[k] + (1 - b1) * gk;
                v[k] = b2 * v[k] + (1 - b2) * gk * gk;
                prm[k] -= (float)(lr * (m[k] / bc1) / (std::sqrt(v[k] / bc2) + eps));
            }
            const int slot_h = (int)((long)step * KQ_TRAIN_HIST / std::max(1, o->steps));
            if (slot_h >= hist_n && hist_n < KQ_TRAIN_HIST)
                R.loss_hist[hist_n++] = lsum / o->batch;
            if (o->verbose && (step % std::max(1, o->steps / 10) == 0 || step == o->steps - 1))
                log_info("[kq] learned: step %d loss %.5f lr %.2e\n", step, lsum / o->batch, lr);
        }
        R.steps = o->steps;
        R.loss_final = mean_loss(w, W, c.stream);
        L.w = w;
        L.have = true;
        R.seconds = secs(t0, clk::now());
        if (rep)
            *rep = R;
        return KUQCD_SUCCESS;
    });
}

void kq_default_refine_options(KqRefineOptions *o)
{
    if (!o)
        return;
    std::memset(o, 0, sizeof *o);
    o->rel_tol = 1e-10;
    o->learned = 1;
    o->mode = KQ_MODE_MIXED_FP32;
    o->net_format = 0;
    o->max_iter = 100000;
    o->delta = 0.1;
    o->chunk = 8;
    o->stall_iter = 300;
    o->verbose = 0;
}

KuQCDStatus kq_fcg_refine(int parity, double sigma, const void *b, void *x, const KqRefineOptions *o,
                          KqRefineReport *rep)
{
    if (!b || !x || !o) {
        set_error("b, x and options must not be nullptr");
        return KUQCD_ERR_NULL_POINTER;
    }
    if (parity != 0 && parity != 1) {
        set_error("parity must be 0 or 1");
        return KUQCD_ERR_INVALID_PARAM;
    }
    if (!(o->rel_tol > 0.0)) {
        set_error("rel_tol must be positive");
        return KUQCD_ERR_INVALID_PARAM;
    }
    return guarded([&] {
        Context &c = ctx();
        const int slot = slot_or_throw();
        const Geom g = slot_geom(slot);
        const size_t n3 = (size_t)3 * g.Vh;
        DevBuf b64(sizeof(double2) * n3), x64(sizeof(double2) * n3), stage;
        upload_vec<double, double>(b, Vec<double>{b64.as<double2>(), g.Vh, n3}, 0, stage, c.stream);
        upload_vec<double, double>(x, Vec<double>{x64.as<double2>(), g.Vh, n3}, 0, stage, c.stream);
        const RefineResult r = fcg_refine_device(slot, parity, sigma, b64.as<double>(), x64.as<double>(), *o);
        download_vec<double, double>(Vec<double>{x64.as<double2>(), g.Vh, n3}, 0, x, stage, c.stream);
        if (rep) {
            std::memset(rep, 0, sizeof *rep);
            rep->iter = r.iter;
            rep->iter_learned = r.iter_learned;
            rep->reliable_updates = r.reliable;
            rep->fallback = r.fallback ? 1 : 0;
            rep->converged = r.converged ? 1 : 0;
            rep->rel_res_initial = r.rel0;
            rep->true_rel_res = r.rel;
            rep->seconds = r.seconds;
        }
        if (!r.converged)
            log_info("WARNING: flexible CG stopped at relative residual %.3e after %d iterations\n", r.rel, r.iter);
        return KUQCD_SUCCESS;
    });
}

}  // extern "C"
