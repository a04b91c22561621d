// A C++ caller of the KuQCD API (include/kuqcd.h), compiled against the MI355X
// library unchanged: it runs the reference's known-answer case
// (test/test_solver.cpp:28-130) -- 4x4x4x1, FP32, every real 1, one shift
// 3.671676313381e-05, ODD_ODD, tol 1e-4, expected 0.007561 -- through the
// drop-in API, reading the process grid from stdin as kuqcd_init does.
#include <cmath>
#include <cstdio>
#include <vector>

#include "kuqcd.h"

int main()
{
    using namespace kuqcd;
    const unsigned lat[4] = {4, 4, 4, 1};
    const size_t vol = 64, volh = 32;
    if (kuqcd_init() != KUQCD_SUCCESS) {
        std::printf("kuqcd_init FAILED\n");
        return 2;
    }
    std::vector<float> fat(vol * 4 * 18, 1.0f), lng(vol * 4 * 18, 1.0f), src(volh * 6, 1.0f), dst(volh * 6, 0.0f);
    KuQCDGaugeParam gp;
    for (int i = 0; i < 4; ++i)
        gp.lat[i] = lat[i];
    gp.halo_depth = 3;
    gp.od = KUQCD_TZYX_ORDER;
    gp.site_subset = KUQCD_ALL;
    gp.gauge_order = KUQCD_NORMAL_ORDER;
    gp.prec = KUQCD_SINGLE;
    gp.recon = KUQCD_GAUGE_RECON_NO;
    gp.link_type = KUQCD_FAT_LINK;
    if (kuqcd_load_gaugefield(fat.data(), &gp) != KUQCD_SUCCESS)
        return 3;
    gp.link_type = KUQCD_LONG_LINK;
    if (kuqcd_load_gaugefield(lng.data(), &gp) != KUQCD_SUCCESS)
        return 3;
    double shifts[1] = {3.671676313381e-05};
    KuQCDInvertParam ip;
    ip.dslash_type = KUQCD_HISQPC_DIRAC;
    ip.inverter_type = KUQCD_CGM;
    ip.matpc_type = KUQCD_ODD_ODD;
    ip.prec = KUQCD_SINGLE;
    ip.num_shift = 1;
    ip.max_iter = 6500;
    ip.iter = 0;
    ip.tol = 1e-4;
    ip.shifts = shifts;
    void *x[1] = {dst.data()};
    if (kuqcd_solve_multishift(x, src.data(), &ip) != KUQCD_SUCCESS)
        return 4;
    double diff = 0;
    for (float v : dst)
        diff += std::fabs(v - 0.007561);
    diff /= (double)dst.size();
    kuqcd_gauge_free();
    kuqcd_end();
    const bool pass = diff < ip.tol;
    std::printf("end test kuqcd_solve_multishift %s (mean |diff| %.3e, iterations %d)\n", pass ? "PASSED" : "FAILED",
                diff, ip.iter);
    return pass ? 0 : 1;
}
