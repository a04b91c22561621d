// Operator-level entry points (hopping term, A = sigma - D D), timing hooks,
// the launch autotuner, synthetic-input generators and host self-tests.
#include <chrono>
#include <cmath>
#include <vector>

#include "kq_gen.h"
#include "kq_solve.h"

namespace kq {

// ---------------------------------------------------------------- dslash on host buffers
template <class LF, class LL, typename VR, typename HR>
static void apply_t(LF fat, LL lng, const Geom &g, int nrhs, const void *src, void *dst, int dst_parity, int nr, int nt,
                    bool mdaggm, double sigma)
{
    Context &c = ctx();
    const size_t n3 = (size_t)3 * g.Vh, nh = (size_t)6 * g.Vh;
    DevBuf s(sizeof(c2_t<VR>) * nrhs * n3), d(sizeof(c2_t<VR>) * nrhs * n3), stage(sizeof(HR) * nh);
    Vec<VR> sv{s.as<c2_t<VR>>(), g.Vh, n3}, dv{d.as<c2_t<VR>>(), g.Vh, n3};
    for (int r = 0; r < nrhs; ++r) {
        KQ_HIP_CHECK(hipMemcpyAsync(stage.as<void>(), static_cast<const HR *>(src) + r * nh, sizeof(HR) * nh,
                                    hipMemcpyDefault, c.stream));
        vec_host_to_soa<HR, VR>(stage.as<HR>(), sv, r, c.stream);
    }
    DslashArgs<LF, LL, VR> a{};
    a.g = g;
    a.fat = fat;
    a.lng = lng;
    a.nrhs = nrhs;
    if (!mdaggm) {
        a.dst_parity = dst_parity;
        a.src = sv;
        a.dst = dv;
        launch_dslash(a, EPI_STORE, nr, nt, c.stream);
    } else {
        DevBuf t(sizeof(c2_t<VR>) * nrhs * n3), sig(sizeof(double) * nrhs);
        DevBuf part(sizeof(double) * nrhs * dslash_nblk(g.Vh, nt));
        std::vector<double> sg(nrhs, sigma);
        KQ_HIP_CHECK(hipMemcpyAsync(sig.as<void>(), sg.data(), sizeof(double) * nrhs, hipMemcpyHostToDevice, c.stream));
        Vec<VR> tv{t.as<c2_t<VR>>(), g.Vh, n3};
        a.dst_parity = 1 - dst_parity;
        a.src = sv;
        a.dst = tv;
        launch_dslash(a, EPI_STORE, nr, nt, c.stream);
        a.dst_parity = dst_parity;
        a.src = tv;
        a.dst = dv;
        a.aux = sv;
        a.sigma = sig.as<double>();
        a.partial = part.as<double>();
        launch_dslash(a, EPI_AXMY_DOT, nr, nt, c.stream);
        KQ_HIP_CHECK(hipStreamSynchronize(c.stream));
    }
    for (int r = 0; r < nrhs; ++r) {
        vec_soa_to_host<VR, HR>(dv, r, stage.as<HR>(), c.stream);
        KQ_HIP_CHECK(hipMemcpyAsync(static_cast<HR *>(dst) + r * nh, stage.as<void>(), sizeof(HR) * nh,
                                    hipMemcpyDefault, c.stream));
        KQ_HIP_CHECK(hipStreamSynchronize(c.stream));
    }
}

static KuQCDStatus apply(int nrhs, const void *src, void *dst, int parity, int mode, int nr, int nt, bool mdaggm,
                         double sigma)
{
    if (src == nullptr || dst == nullptr) {
        set_error("src or dst is nullptr");
        return KUQCD_ERR_NULL_POINTER;
    }
    if (parity != 0 && parity != 1) {
        set_error("parity must be 0 (even) or 1 (odd)");
        return KUQCD_ERR_INVALID_PARAM;
    }
    const int slot = active_slot();
    if (slot < 0) {
        set_error("no init gauge_smeared and gauge_naik");
        return KUQCD_ERR_UNKNOWN_TYPE;
    }
    if (nrhs <= 0) {
        set_error("nrhs must be positive");
        return KUQCD_ERR_INVALID_PARAM;
    }
    if (nr != 1 && nr != 2 && nr != 3 && nr != 4 && nr != 8) {
        set_error("nr must be 1, 2, 3, 4 or 8");
        return KUQCD_ERR_INVALID_PARAM;
    }
    if (nt != 64 && nt != 128 && nt != 256) {
        set_error("nt must be 64, 128 or 256");
        return KUQCD_ERR_INVALID_PARAM;
    }
    Context &c = ctx();
    const Geom g = slot_geom(slot);
    const bool dbl = c.fat[slot].prec == KUQCD_DOUBLE;
    if (mode == KQ_MODE_REFERENCE || mode == KQ_MODE_FP64) {
        if (dbl)
            with_fp64_links(slot, [&](auto fat, auto lng) {
                apply_t<decltype(fat), decltype(lng), double, double>(fat, lng, g, nrhs, src, dst, parity, nr, nt,
                                                                      mdaggm, sigma);
            });
        else if (mode == KQ_MODE_REFERENCE)
            with_fp32_links(slot, [&](auto fat, auto lng) {
                apply_t<decltype(fat), decltype(lng), float, float>(fat, lng, g, nrhs, src, dst, parity, nr, nt, mdaggm,
                                                                    sigma);
            });
        else {
            set_error("FP64 mode needs a double-precision gauge field");
            return KUQCD_ERR_UNSUPPORTED;
        }
        return KUQCD_SUCCESS;
    }
    if (!dbl) {
        set_error("mixed modes need a double-precision gauge field");
        return KUQCD_ERR_UNSUPPORTED;
    }
    if (is_mixed_mode(mode)) {
        with_mixed_links(slot, mode, [&](auto fat, auto lng) {
            apply_t<decltype(fat), decltype(lng), float, double>(fat, lng, g, nrhs, src, dst, parity, nr, nt, mdaggm,
                                                                 sigma);
        });
        return KUQCD_SUCCESS;
    }
    set_error("unknown mode");
    return KUQCD_ERR_INVALID_PARAM;
}

// ---------------------------------------------------------------- timing
template <class LF, class LL, typename VR>
static double time_dslash(LF fat, LL lng, const Geom &g, int nrhs, int nr, int nt, int iters, int cap)
{
    Context &c = ctx();
    const size_t n3 = (size_t)3 * g.Vh;
    DevBuf s(sizeof(c2_t<VR>) * nrhs * n3), d(sizeof(c2_t<VR>) * nrhs * n3);
    KQ_HIP_CHECK(hipMemsetAsync(s.as<void>(), 0, s.bytes(), c.stream));
    DslashArgs<LF, LL, VR> a{};
    a.g = g;
    a.fat = fat;
    a.lng = lng;
    a.nrhs = nrhs;
    a.dst_parity = 0;
    a.src = Vec<VR>{s.as<c2_t<VR>>(), g.Vh, n3};
    a.dst = Vec<VR>{d.as<c2_t<VR>>(), g.Vh, n3};
    for (int i = 0; i < 2; ++i)
        launch_dslash(a, EPI_STORE, nr, nt, c.stream, cap);
    hipEvent_t e0, e1;
    KQ_HIP_CHECK(hipEventCreate(&e0));
    KQ_HIP_CHECK(hipEventCreate(&e1));
    KQ_HIP_CHECK(hipEventRecord(e0, c.stream));
    for (int i = 0; i < iters; ++i) {
        a.dst_parity = i & 1;
        launch_dslash(a, EPI_STORE, nr, nt, c.stream, cap);
    }
    KQ_HIP_CHECK(hipEventRecord(e1, c.stream));
    KQ_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    KQ_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return (double)ms / iters;
}

// Algorithmic bytes per output site of a hopping term with link formats LF, LL and vectors of
// real type R: every link of the lattice once (8 fat + 8 long per site), one source and one
// result vector per right-hand side
template <class LF, class LL, typename R> static double site_bytes(int nrhs)
{
    return 8.0 * (LF::kBytes + LL::kBytes) + nrhs * 2.0 * 3 * sizeof(c2_t<R>);
}

// dispatch a timing run of the format `mode` on slot; *bsite: its algorithmic bytes per site
static double time_mode(int slot, int mode, int nrhs, int nr, int nt, int iters, double *bsite, int cap = 0)
{
    Context &c = ctx();
    const Geom g = slot_geom(slot);
    const bool dbl = c.fat[slot].prec == KUQCD_DOUBLE;
    double ms = 0.0;
    if (mode == KQ_MODE_REFERENCE || mode == KQ_MODE_FP64) {
        if (dbl)
            with_fp64_links(slot, [&](auto fat, auto lng) {
                using LF = decltype(fat);
                using LL = decltype(lng);
                ms = time_dslash<LF, LL, double>(fat, lng, g, nrhs, nr, nt, iters, cap);
                *bsite = site_bytes<LF, LL, double>(nrhs);
            });
        else
            with_fp32_links(slot, [&](auto fat, auto lng) {
                using LF = decltype(fat);
                using LL = decltype(lng);
                ms = time_dslash<LF, LL, float>(fat, lng, g, nrhs, nr, nt, iters, cap);
                *bsite = site_bytes<LF, LL, float>(nrhs);
            });
        return ms;
    }
    with_mixed_links(slot, mode, [&](auto fat, auto lng) {
        using LF = decltype(fat);
        using LL = decltype(lng);
        ms = time_dslash<LF, LL, float>(fat, lng, g, nrhs, nr, nt, iters, cap);
        *bsite = site_bytes<LF, LL, float>(nrhs);
    });
    return ms;
}

std::pair<int, int> tuned_launch(int slot, int mode, int nrhs)
{
    Context &c = ctx();
    const Geom g = slot_geom(slot);
    if (mode == KQ_MODE_REFERENCE && c.fat[slot].prec == KUQCD_DOUBLE)
        mode = KQ_MODE_FP64;
    const auto key = std::make_tuple(mode, nrhs, g.Vh);
    auto it = c.tune.find(key);
    if (it != c.tune.end())
        return it->second;
    // Each blocking NR is timed once, on a single group of NR right-hand sides, and a launch
    // for nrhs systems is priced as ceil(nrhs / NR) such groups (the groups are independent
    // workgroups of the same kernel). Blockings of 1, 2, 3, 4 and 8 are tried below 2 * nrhs:
    // the kernel pads a partial block with the last live right-hand side and writes nothing
    // for it (three live systems in a block of four cost the padded fourth's arithmetic). FP64 stops at NR = 2 (register pressure). Each (NR, block size) is timed
    // without an occupancy cap and with caps of 5..2 waves per SIMD; the best cap is kept
    // (tuned_cap).
    static const int kNr[5] = {1, 2, 3, 4, 8}, kNt[3] = {64, 128, 256}, kCap[5] = {0, 5, 4, 3, 2};
    auto &tg = c.tune_group[{mode, g.Vh}];
    if (tg[0] == 0.0)
        tg.fill(-1.0);
    double best = 1e30;
    std::pair<int, int> pick{1, 256};
    double bsite = 0.0;
    for (int i = 0; i < 5; ++i) {
        const int nr = kNr[i];
        if (nr > 1 && (nr >= 2 * nrhs || (mode == KQ_MODE_FP64 && nr > 2)))
            continue;
        for (int j = 0; j < 3; ++j) {
            double &ms = tg[3 * i + j];
            if (ms < 0) {
                int best_cap = 0;
                for (int cap : kCap) {
                    const double t = time_mode(slot, mode, nr, nr, kNt[j], 8, &bsite, cap);
                    if (ms < 0 || t < ms) {
                        ms = t;
                        best_cap = cap;
                    }
                }
                c.tune_cap[std::make_tuple(mode, g.Vh, nr, kNt[j])] = best_cap;
            }
            const double cost = ms * ((nrhs + nr - 1) / nr);
            if (cost < best) {
                best = cost;
                pick = {nr, kNt[j]};
            }
        }
    }
    c.tune[key] = pick;
    if (c.verbosity > 1)
        log_info("autotune mode %d nrhs %d: nr %d nt %d cap %d (%.3f ms)\n", mode, nrhs, pick.first, pick.second,
                 tuned_cap(slot, mode, pick.first, pick.second), best);
    return pick;
}

int tuned_cap(int slot, int mode, int nr, int nt)
{
    Context &c = ctx();
    const Geom g = slot_geom(slot);
    if (mode == KQ_MODE_REFERENCE && c.fat[slot].prec == KUQCD_DOUBLE)
        mode = KQ_MODE_FP64;
    const auto key = std::make_tuple(mode, g.Vh, nr, nt);
    auto it = c.tune_cap.find(key);
    if (it == c.tune_cap.end()) {
        (void)tuned_launch(slot, mode, nr);  // times every blocking up to nr
        it = c.tune_cap.find(key);
    }
    return it == c.tune_cap.end() ? 0 : it->second;
}

}  // namespace kq

using namespace kq;

extern "C" {

KuQCDStatus kq_dslash(const void *src, void *dst, int dst_parity, int mode)
{
    try {
        return apply(1, src, dst, dst_parity, mode, 1, 256, false, 0.0);
    } catch (const std::bad_alloc &) {
        set_error("out of memory");
        return KUQCD_ERR_NO_MEMORY;
    } catch (const std::exception &e) {
        set_error(e.what());
        return KUQCD_ERR_UNKNOWN_ERROR;
    }
}

KuQCDStatus kq_dslash_batched(int nrhs, const void *src, void *dst, int dst_parity, int mode, int nr, int nt)
{
    try {
        return apply(nrhs, src, dst, dst_parity, mode, nr, nt, false, 0.0);
    } catch (const std::bad_alloc &) {
        set_error("out of memory");
        return KUQCD_ERR_NO_MEMORY;
    } catch (const std::exception &e) {
        set_error(e.what());
        return KUQCD_ERR_UNKNOWN_ERROR;
    }
}

KuQCDStatus kq_mdaggm(const void *src, void *dst, int parity, double sigma, int mode)
{
    try {
        return apply(1, src, dst, parity, mode, 1, 256, true, sigma);
    } catch (const std::bad_alloc &) {
        set_error("out of memory");
        return KUQCD_ERR_NO_MEMORY;
    } catch (const std::exception &e) {
        set_error(e.what());
        return KUQCD_ERR_UNKNOWN_ERROR;
    }
}

KuQCDStatus kq_bench_dslash(int mode, int nrhs, int nr, int nt, int iters, KqBenchResult *out)
{
    try {
        if (out == nullptr) {
            set_error("out is nullptr");
            return KUQCD_ERR_NULL_POINTER;
        }
        const int slot = active_slot();
        if (slot < 0) {
            set_error("no init gauge_smeared and gauge_naik");
            return KUQCD_ERR_UNKNOWN_TYPE;
        }
        if (nrhs <= 0 || iters <= 0) {
            set_error("nrhs and iters must be positive");
            return KUQCD_ERR_INVALID_PARAM;
        }
        int cap = 0;
        if (nr <= 0 || nt <= 0) {
            const auto p = tuned_launch(slot, mode, nrhs);
            nr = nr > 0 ? nr : p.first;
            nt = nt > 0 ? nt : p.second;
            cap = tuned_cap(slot, mode, nr, nt);
        }
        double bsite = 0.0;
        const double ms = time_mode(slot, mode, nrhs, nr, nt, iters, &bsite, cap);
        const Geom g = slot_geom(slot);
        out->ms = ms;
        out->bytes = bsite * (double)g.Vh;
        out->gbs = out->bytes / (ms * 1e-3) / 1e9;
        out->gflops = 1146.0 * (double)g.Vh * nrhs / (ms * 1e-3) / 1e9;
        out->hbm_frac = out->gbs / 8000.0;
        out->nr = nr;
        out->nt = nt;
        return KUQCD_SUCCESS;
    } catch (const std::bad_alloc &) {
        set_error("out of memory");
        return KUQCD_ERR_NO_MEMORY;
    } catch (const std::exception &e) {
        set_error(e.what());
        return KUQCD_ERR_UNKNOWN_ERROR;
    }
}

KuQCDStatus kq_bench_copy(size_t bytes, int iters, KqBenchResult *out)
{
    try {
        if (out == nullptr) {
            set_error("out is nullptr");
            return KUQCD_ERR_NULL_POINTER;
        }
        ensure_device();
        Context &c = ctx();
        bytes = (bytes + 15) / 16 * 16;
        DevBuf a(bytes), b(bytes);
        KQ_HIP_CHECK(hipMemsetAsync(a.as<void>(), 0, bytes, c.stream));
        device_copy(a.as<void>(), b.as<void>(), bytes, c.stream);
        hipEvent_t e0, e1;
        KQ_HIP_CHECK(hipEventCreate(&e0));
        KQ_HIP_CHECK(hipEventCreate(&e1));
        KQ_HIP_CHECK(hipEventRecord(e0, c.stream));
        for (int i = 0; i < iters; ++i)
            device_copy(a.as<void>(), b.as<void>(), bytes, c.stream);
        KQ_HIP_CHECK(hipEventRecord(e1, c.stream));
        KQ_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        KQ_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        out->ms = ms / iters;
        out->bytes = 2.0 * (double)bytes;
        out->gbs = out->bytes / (out->ms * 1e-3) / 1e9;
        out->gflops = 0;
        out->hbm_frac = out->gbs / 8000.0;
        out->nr = out->nt = 0;
        return KUQCD_SUCCESS;
    } catch (const std::exception &e) {
        set_error(e.what());
        return KUQCD_ERR_UNKNOWN_ERROR;
    }
}

KuQCDStatus kq_autotune_dslash(int mode, int nrhs, int *nr, int *nt)
{
    try {
        const int slot = active_slot();
        if (slot < 0) {
            set_error("no init gauge_smeared and gauge_naik");
            return KUQCD_ERR_UNKNOWN_TYPE;
        }
        const auto p = tuned_launch(slot, mode, nrhs);
        if (nr)
            *nr = p.first;
        if (nt)
            *nt = p.second;
        return KUQCD_SUCCESS;
    } catch (const std::exception &e) {
        set_error(e.what());
        return KUQCD_ERR_UNKNOWN_ERROR;
    }
}

int kq_gen_gauge(double *fat, double *lng, const unsigned lat[4], int kind, double eps, double fat_scale,
                 double c_naik, uint64_t seed_fat, uint64_t seed_long, int phases)
{
    if (lat == nullptr)
        return -1;
    return gen_gauge<double>(fat, lng, lat, kind, eps, fat_scale, c_naik, seed_fat, seed_long, phases);
}

int kq_gen_source(double *b, const unsigned lat[4], int parity, uint64_t seed)
{
    if (b == nullptr || lat == nullptr)
        return -1;
    return gen_source<double>(b, lat, parity, seed);
}

long kq_selftest_fastdiv(uint32_t d, uint32_t n_max)
{
    if (d == 0)
        return -1;
    FastDiv f(d);
    long bad = 0;
    for (uint32_t n = 0; n < n_max; ++n)
        if (f.div(n) != n / d)
            ++bad;
    return bad;
}

int kq_selftest_fold_keeps(int discard, int k, int max_term)
{
    return fold_keeps(discard, k, max_term) ? 1 : 0;
}

void kq_selftest_seg_coef(const double *h, int J, int k, int keep_x, double *out)
{
    if (h == nullptr || out == nullptr || k < 0 || k >= kMaxShift)
        return;
    for (int m = -1; m < kHistMax; ++m)
        seg_coefficient(h, J, k, m, keep_x != 0, out[1 + m], out[kHistMax + 2 + m]);
}

long kq_neighbour_index(const int lat[4], int q, long s, int mu, int h)
{
    const Geom g = make_geom(lat);
    const SiteCoord sc = site_coords(g, (uint32_t)s, q);
    switch (mu) {
    case 0:
        return nb_index<0>(g, (uint32_t)s, sc, h);
    case 1:
        return nb_index<1>(g, (uint32_t)s, sc, h);
    case 2:
        return nb_index<2>(g, (uint32_t)s, sc, h);
    default:
        return nb_index<3>(g, (uint32_t)s, sc, h);
    }
}

}  // extern "C"
