// The KuQCD API (kuqcd.h:29-34) and its C-ABI forms, on the MI355X.
//
// Validation follows the reference's order and status codes
// (kuqcd.cpp:41-101, 171-212, 288-445). Where the reference has undefined
// behaviour the library returns an error instead:
//   halo_depth not in {0,3}, gauge site_subset != ALL, odd Lx  -> INVALID_PARAM (kuqcd.cpp:62-99, 342)
//   param->prec != size of the loaded gauge's real            -> INVALID_PARAM (lattice_field.cpp:31)
//   KUQCD_HISQ_DIRAC (non even-odd) solve: the reference gives it half-lattice fields that its
//     operator indexes as full ones (dirac_hisq.cpp:130, 422). kq_solve_multishift solves on
//     fields of both parities (V sites, even block first); the extensions -> UNSUPPORTED
//   more than 32 shifts                                       -> UNSUPPORTED
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "kq_context.h"
#include "kq_learned.h"
#include "kq_solve.h"
#include "kuqcd.h"
#include "kuqcd_hip.h"

namespace kq {

Context &ctx()
{
    static Context c;
    return c;
}

void ensure_device()
{
    Context &c = ctx();
    if (c.device_ready)
        return;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        throw std::runtime_error("no HIP device is visible");
    KQ_HIP_CHECK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    KQ_HIP_CHECK(hipStreamCreateWithFlags(&c.copy_stream, hipStreamNonBlocking));
    DevPool::get().set_stream(c.stream);
    (void)noalloc_scratch();  // the no-allocation entry point's scratch address on this device
    c.device_ready = true;
}

void log_info(const char *fmt, ...)
{
    if (ctx().verbosity <= 0)
        return;
    va_list ap;
    va_start(ap, fmt);
    vprintf(fmt, ap);
    va_end(ap);
    fflush(stdout);
}

void set_error(const std::string &msg)
{
    ctx().last_error = msg;
    if (ctx().verbosity > 0)
        fprintf(stderr, "[kuqcd] error: %s\n", msg.c_str());
}

int active_slot()
{
    Context &c = ctx();
    for (int i = 0; i < 4; ++i)
        if (c.fat[i].loaded && c.lng[i].loaded)
            return i;
    return -1;
}

Geom slot_geom(int slot)
{
    int L[4];
    for (int i = 0; i < 4; ++i)
        L[i] = (int)ctx().fat[slot].lat[i];
    return make_geom(L);
}

LinkD link_d(const GaugeSlot &s)
{
    const uint32_t Vh = (uint32_t)((uint64_t)s.lat[0] * s.lat[1] * s.lat[2] * s.lat[3] / 2);
    return LinkD{s.kf.as<char>(), nb64_of(Vh)};
}
LinkD14 link_d14(const GaugeSlot &s)
{
    const uint32_t Vh = (uint32_t)((uint64_t)s.lat[0] * s.lat[1] * s.lat[2] * s.lat[3] / 2);
    return LinkD14{s.kfr.as<char>(), nb64_of(Vh)};
}


LinkF link_f(const GaugeSlot &s)
{
    const uint32_t Vh = (uint32_t)((uint64_t)s.lat[0] * s.lat[1] * s.lat[2] * s.lat[3] / 2);
    return LinkF{s.kf.as<char>(), nb64_of(Vh)};
}


// run a packer that counts the links without its form; true when none is missing it
template <typename C2>
static bool all_links_pass(void (*pack)(const C2 *, char *, uint32_t, int *, hipStream_t), const C2 *canon, DevBuf &dst,
                           uint32_t Vh)
{
    Context &c = ctx();
    DevBuf fail(sizeof(int));
    KQ_HIP_CHECK(hipMemsetAsync(fail.as<void>(), 0, sizeof(int), c.stream));
    pack(canon, dst.as<char>(), Vh, fail.as<int>(), c.stream);
    int nfail = 0;
    KQ_HIP_CHECK(hipMemcpyAsync(&nfail, fail.as<void>(), sizeof(int), hipMemcpyDeviceToHost, c.stream));
    KQ_HIP_CHECK(hipStreamSynchronize(c.stream));
    return nfail == 0;
}

DerivedFormats &derived(int slot, bool want_f32, bool want_f16)
{
    Context &c = ctx();
    DerivedFormats &d = c.der[slot];
    const GaugeSlot &f = c.fat[slot], &l = c.lng[slot];
    if (f.prec != KUQCD_DOUBLE || l.prec != KUQCD_DOUBLE)
        throw std::invalid_argument("lower-precision formats are derived from a double-precision gauge");
    const uint32_t Vh = slot_geom(slot).Vh;
    if (want_f32 && !d.f32) {
        d.fat_f.alloc(link_field_bytes<LinkF>(Vh));
        d.lng_f.alloc(link_field_bytes<LinkF>(Vh));
        canon_to_linkF<double>(f.soa.as<double2>(), d.fat_f.as<char>(), Vh, c.stream);
        canon_to_linkF<double>(l.soa.as<double2>(), d.lng_f.as<char>(), Vh, c.stream);
        // long links also as 12 reals when every one is b * SU(3), or w * SU(3) with w complex
        d.lng_f12.alloc(link_field_bytes<LinkF12>(Vh));
        if (all_links_pass(canon_to_linkF12, l.soa.as<double2>(), d.lng_f12, Vh)) {
            d.f_long = kLong12;
        } else {
            d.lng_f12.alloc(link_field_bytes<LinkF12C>(Vh));
            if (all_links_pass(canon_to_linkF12C, l.soa.as<double2>(), d.lng_f12, Vh))
                d.f_long = kLong12C;
            else
                d.lng_f12.release();
        }
        d.f32 = true;
    }
    if (want_f16 && !d.f16) {
        d.fat_h.alloc(link_field_bytes<LinkH>(Vh));
        canon_to_linkH(f.soa.as<double2>(), d.fat_h.as<char>(), Vh, c.stream);
        // long links: 12 reals with a real scale, else with a complex scale, else 18 reals
        d.lng_h.alloc(link_field_bytes<LinkH12>(Vh));
        if (all_links_pass(canon_to_linkH12, l.soa.as<double2>(), d.lng_h, Vh)) {
            d.h_long = kLong12;
        } else {
            d.lng_h.alloc(link_field_bytes<LinkH12C>(Vh));
            if (all_links_pass(canon_to_linkH12C, l.soa.as<double2>(), d.lng_h, Vh)) {
                d.h_long = kLong12C;
            } else {
                d.lng_h.alloc(link_field_bytes<LinkH>(Vh));
                canon_to_linkH(l.soa.as<double2>(), d.lng_h.as<char>(), Vh, c.stream);
                KQ_HIP_CHECK(hipStreamSynchronize(c.stream));
                d.h_long = kLong18;
            }
        }
        d.f16 = true;
    }
    return d;
}

static void free_derived(int slot)
{
    ctx().der[slot] = DerivedFormats();
}

// ---------------------------------------------------------------- init
static KuQCDStatus init_with_grid(const int grid[4])
{
    Context &c = ctx();
    for (int i = 0; i < 4; ++i)
        c.grid[i] = grid[i];
    // one process per GPU: the reference requires prod(grid) == number of ranks (kuqcd.cpp:296-299)
    const long prod = (long)grid[0] * grid[1] * grid[2] * grid[3];
    if (prod != 1) {
        set_error("g_grid_dim is not equal np");
        return KUQCD_ERR_INVALID_PARAM;
    }
    return KUQCD_SUCCESS;
}

static std::string trim(const std::string &s)
{
    const size_t a = s.find_first_not_of(" \t\r");
    if (a == std::string::npos)
        return "";
    const size_t b = s.find_last_not_of(" \t\r");
    return s.substr(a, b - a + 1);
}

// gridx=.. lines from stdin, as CmdLineParam::ReadParam (command_line.cpp:31-84);
// keys not given stay 1.
static KuQCDStatus init_from_stdin()
{
    int grid[4] = {1, 1, 1, 1};
    const char *keys[4] = {"gridx", "gridy", "gridz", "gridt"};
    std::string line;
    while (std::getline(std::cin, line)) {
        const size_t eq = line.find('=');
        if (eq == std::string::npos) {
            std::cerr << "ignore invalid lines (no =): " << line << std::endl;
            continue;
        }
        const std::string key = trim(line.substr(0, eq)), val = trim(line.substr(eq + 1));
        int which = -1;
        for (int i = 0; i < 4; ++i)
            if (key == keys[i])
                which = i;
        if (which < 0) {
            std::cerr << "unknown variable" << key << "'" << std::endl;
            continue;
        }
        try {
            grid[which] = std::stoi(val);
        } catch (const std::exception &) {
            std::cerr << "invalid number '" << val << "' corresponding variable '" << key << "'" << std::endl;
        }
    }
    if (ctx().verbosity > 0)
        std::cout << "gridx = " << grid[0] << "  gridy = " << grid[1] << "  gridz = " << grid[2]
                  << "  gridt = " << grid[3] << std::endl;
    return init_with_grid(grid);
}

// ---------------------------------------------------------------- load

template <typename HR>
static void upload_links(const HR *h, GaugeSlot &slot, double scale, int prec, bool long_links)
{
    Context &c = ctx();
    const uint64_t V = (uint64_t)slot.lat[0] * slot.lat[1] * slot.lat[2] * slot.lat[3];
    const uint32_t Vh = (uint32_t)(V / 2);
    const size_t ncomp = (size_t)V * 36;
    slot.soa.alloc((prec == KUQCD_DOUBLE ? sizeof(double2) : sizeof(float2)) * ncomp);
    // stream the host array through a device staging buffer in chunks of sites
    const size_t chunk = std::min<uint64_t>(V, (uint64_t)1 << 20);
    DevBuf stage(sizeof(HR) * 72 * chunk);
    for (uint64_t i0 = 0; i0 < V; i0 += chunk) {
        const size_t n = (size_t)std::min<uint64_t>(chunk, V - i0);
        copy_from_host(stage.as<void>(), h + 72 * i0, sizeof(HR) * 72 * n, c.stream);
        if (prec == KUQCD_DOUBLE)
            gauge_host_to_soa<HR, double>(stage.as<HR>(), i0, n, slot.soa.as<double2>(), Vh, scale, c.stream);
        else
            gauge_host_to_soa<HR, float>(stage.as<HR>(), i0, n, slot.soa.as<float2>(), Vh, scale, c.stream);
        KQ_HIP_CHECK(hipStreamSynchronize(c.stream));
    }
    // the kernel format of the loaded precision
    if (prec == KUQCD_DOUBLE) {
        slot.kf.alloc(link_field_bytes<LinkD>(Vh));
        canon_to_linkD(slot.soa.as<double2>(), slot.kf.as<char>(), Vh, c.stream);
    } else {
        slot.kf.alloc(link_field_bytes<LinkF>(Vh));
        canon_to_linkF<float>(slot.soa.as<float2>(), slot.kf.as<char>(), Vh, c.stream);
    }
    // long links also in fewer reals when every one has that form: 14 in double (a complex
    // multiple of an SU(3) matrix, to 1e-13); 12 in float, with a real or else a complex scale
    // (to 1e-6)
    if (long_links) {
        slot.lform = kLong18;
        if (prec == KUQCD_DOUBLE) {
            slot.kfr.alloc(link_field_bytes<LinkD14>(Vh));
            if (all_links_pass(canon_to_linkD14, slot.soa.as<double2>(), slot.kfr, Vh))
                slot.lform = kLong12C;
        } else {
            slot.kfr.alloc(link_field_bytes<LinkF12>(Vh));
            if (all_links_pass(canon_to_linkF12, slot.soa.as<float2>(), slot.kfr, Vh)) {
                slot.lform = kLong12;
            } else {
                slot.kfr.alloc(link_field_bytes<LinkF12C>(Vh));
                if (all_links_pass(canon_to_linkF12C, slot.soa.as<float2>(), slot.kfr, Vh))
                    slot.lform = kLong12C;
            }
        }
        if (slot.lform == kLong18)
            slot.kfr.release();
    }
    KQ_HIP_CHECK(hipStreamSynchronize(c.stream));
}

static KuQCDStatus load_gauge(void *h_gauge, KuQCDGaugeParam *param)
{
    // kuqcd.cpp:306-332
    if (h_gauge == nullptr || param == nullptr) {
        set_error("h_gauge or param is nullptr");
        return KUQCD_ERR_NULL_POINTER;
    }
    if (param->gauge_order != KUQCD_BATCH_ORDER && param->gauge_order != KUQCD_NORMAL_ORDER) {
        set_error("gauge_order is invalid");
        return KUQCD_ERR_INVALID_PARAM;
    }
    if (param->link_type != KUQCD_FAT_LINK && param->link_type != KUQCD_LONG_LINK) {
        set_error("link_type is invalid");
        return KUQCD_ERR_INVALID_PARAM;
    }
    if (param->recon != KUQCD_GAUGE_RECON_NO) {
        if (param->recon == KUQCD_GAUGE_RECON_U14 || param->recon == KUQCD_GAUGE_RECON_12 ||
            param->recon == KUQCD_GAUGE_RECON_8) {
            set_error("recon is not support");
            return KUQCD_ERR_UNSUPPORTED;
        }
        set_error("recon is invalid");
        return KUQCD_ERR_INVALID_PARAM;
    }
    if (param->lat[0] == 0 || param->lat[1] == 0 || param->lat[2] == 0 || param->lat[3] == 0) {
        set_error("param->lat is 0");
        return KUQCD_ERR_INVALID_PARAM;
    }
    // set_latbase checks (kuqcd.cpp:45-61)
    if (param->prec != KUQCD_SINGLE && param->prec != KUQCD_DOUBLE) {
        if (param->prec == KUQCD_HALF) {
            set_error("prec is not support");
            return KUQCD_ERR_UNSUPPORTED;
        }
        set_error("prec is not valid");
        return KUQCD_ERR_INVALID_PARAM;
    }
    if (param->site_subset < KUQCD_EVEN || param->site_subset >= KUQCD_INVALID_PARITY) {
        set_error("site_subset is not valid");
        return KUQCD_ERR_INVALID_PARAM;
    }
    // cases the reference leaves undefined
    if (param->halo_depth != 0 && param->halo_depth != 3) {
        set_error("halo_depth must be 0 or 3");
        return KUQCD_ERR_INVALID_PARAM;
    }
    if (param->site_subset != KUQCD_ALL) {
        set_error("the gauge field must hold both parities (site_subset = KUQCD_ALL)");
        return KUQCD_ERR_INVALID_PARAM;
    }
    if (param->lat[0] % 2 != 0) {
        set_error("x direction lattice should be even");
        return KUQCD_ERR_INVALID_PARAM;
    }
    for (int i = 0; i < 4; ++i)
        if (param->lat[i] > 65535) {
            set_error("lattice extents are limited to 65535");
            return KUQCD_ERR_INVALID_PARAM;
        }
    const uint64_t V = (uint64_t)param->lat[0] * param->lat[1] * param->lat[2] * param->lat[3];
    if (V / 2 >= ((uint64_t)1 << 31)) {
        set_error("lattice too large (half volume must be below 2^31 sites)");
        return KUQCD_ERR_UNSUPPORTED;
    }
    ensure_device();
    Context &c = ctx();
    const int slot = slot_index(param->gauge_order, param->prec);
    GaugeSlot &gs = (param->link_type == KUQCD_FAT_LINK) ? c.fat[slot] : c.lng[slot];
    gs.loaded = false;
    gs.soa.release();
    gs.kf.release();
    gs.kfr.release();
    gs.lform = kLong18;
    free_derived(slot);
    for (int i = 0; i < 4; ++i)
        gs.lat[i] = param->lat[i];
    gs.prec = param->prec;
    const double scale = (param->link_type == KUQCD_FAT_LINK) ? 0.5 : (-1.0 / 48.0);
    if (param->prec == KUQCD_DOUBLE)
        upload_links<double>(static_cast<const double *>(h_gauge), gs, scale, KUQCD_DOUBLE,
                             param->link_type == KUQCD_LONG_LINK);
    else
        upload_links<float>(static_cast<const float *>(h_gauge), gs, scale, KUQCD_SINGLE,
                            param->link_type == KUQCD_LONG_LINK);
    gs.loaded = true;
    return KUQCD_SUCCESS;
}

// ---------------------------------------------------------------- solve
// Checks of kuqcd_solve_multishift + kuqcd_inverter (kuqcd.cpp:407-414, 176-199, 214)
// and of the extensions; on success *slot is the gauge pair to use. full_ok: the caller
// solves KUQCD_HISQ_DIRAC on fields of both parities.
KuQCDStatus check_solve(void **x, void *b, KuQCDInvertParam *param, int nrhs, int *slot, bool full_ok = false)
{
    if (x == nullptr || b == nullptr || param == nullptr) {
        set_error("hp_x or hp_b or param is nullptr");
        return KUQCD_ERR_NULL_POINTER;
    }
    if (param->matpc_type != KUQCD_EVEN_EVEN && param->matpc_type != KUQCD_ODD_ODD) {
        set_error("matpc_type is error!");
        return KUQCD_ERR_INVALID_PARAM;
    }
    const int s = active_slot();
    if (s < 0) {
        set_error("no init gauge_smeared and gauge_naik");
        return KUQCD_ERR_UNKNOWN_TYPE;
    }
    if (param->inverter_type != KUQCD_CGM) {
        set_error("Inverter is not support");
        return KUQCD_ERR_UNSUPPORTED;
    }
    if (param->dslash_type == KUQCD_WILSON_DIRAC || param->dslash_type == KUQCD_WILSONPC_DIRAC) {
        set_error("dirac type should be hisq or hisq pc!");
        return KUQCD_ERR_UNSUPPORTED;
    }
    if (param->dslash_type != KUQCD_HISQ_DIRAC && param->dslash_type != KUQCD_HISQPC_DIRAC) {
        set_error("dirac type is error!");
        return KUQCD_ERR_INVALID_PARAM;
    }
    if (param->num_shift <= 0) {
        set_error("num_shift < 0!");
        return KUQCD_ERR_INVALID_PARAM;
    }
    if (param->prec != KUQCD_SINGLE && param->prec != KUQCD_DOUBLE) {
        if (param->prec == KUQCD_HALF) {
            set_error("prec is not support");
            return KUQCD_ERR_UNSUPPORTED;
        }
        set_error("prec is not valid");
        return KUQCD_ERR_INVALID_PARAM;
    }
    const Context &c = ctx();
    if ((int)param->prec != c.fat[s].prec) {
        set_error("param->prec must equal the precision of the loaded gauge field");
        return KUQCD_ERR_INVALID_PARAM;
    }
    if (param->dslash_type == KUQCD_HISQ_DIRAC && !full_ok) {
        set_error("this entry point solves the even-odd preconditioned operator (KUQCD_HISQPC_DIRAC) only");
        return KUQCD_ERR_UNSUPPORTED;
    }
    if (param->num_shift > kMaxShift) {
        set_error("at most 32 shifts are supported");
        return KUQCD_ERR_UNSUPPORTED;
    }
    if (param->shifts == nullptr) {
        set_error("param->shifts is nullptr");
        return KUQCD_ERR_NULL_POINTER;
    }
    for (int i = 0; i < nrhs * param->num_shift; ++i)
        if (x[i] == nullptr) {
            set_error("hp_x[k] is nullptr");
            return KUQCD_ERR_NULL_POINTER;
        }
    for (int i = 0; i < 4; ++i)
        if (c.fat[s].lat[i] != c.lng[s].lat[i]) {
            set_error("fat and long links were loaded with different lattices");
            return KUQCD_ERR_INVALID_PARAM;
        }
    *slot = s;
    return KUQCD_SUCCESS;
}

template <typename Fn> static KuQCDStatus guarded(Fn fn)
{
    try {
        return fn();
    } catch (const std::bad_alloc &) {
        set_error("out of memory");
        return KUQCD_ERR_NO_MEMORY;
    } catch (const std::exception &e) {
        set_error(e.what());
        return KUQCD_ERR_UNKNOWN_ERROR;
    } catch (...) {
        set_error("unknown exception");
        return KUQCD_ERR_UNKNOWN_ERROR;
    }
}

static void free_all_gauges()
{
    Context &c = ctx();
    if (!c.device_ready)
        return;
    (void)hipStreamSynchronize(c.stream);
    for (int i = 0; i < 4; ++i) {
        c.fat[i] = GaugeSlot();
        c.lng[i] = GaugeSlot();
        c.der[i] = DerivedFormats();
    }
    c.tune.clear();
    c.tune_group.clear();
    c.tune_cap.clear();
    DevPool::get().trim();
}

}  // namespace kq

using namespace kq;

// ================================================================ C ABI
extern "C" {

KuQCDStatus kq_init(void)
{
    return guarded([] { return init_from_stdin(); });
}
KuQCDStatus kq_init_grid(const int grid[4])
{
    if (grid == nullptr) {
        set_error("grid is nullptr");
        return KUQCD_ERR_NULL_POINTER;
    }
    return guarded([&] { return init_with_grid(grid); });
}
KuQCDStatus kq_load_gaugefield(void *h_gauge, KuQCDGaugeParam *param)
{
    return guarded([&] { return load_gauge(h_gauge, param); });
}
KuQCDStatus kq_solve_multishift(void **x, void *b, KuQCDInvertParam *param)
{
    return guarded([&] {
        int slot = -1;
        KuQCDStatus st = check_solve(x, b, param, 1, &slot, true);
        if (st != KUQCD_SUCCESS)
            return st;
        return solve_reference(slot, x, b, *param);
    });
}
void kq_gauge_free(void)
{
    try {
        free_all_gauges();
    } catch (...) {
    }
}
void kq_end(void)
{
    try {
        Context &c = ctx();
        if (c.device_ready) {
            (void)hipStreamSynchronize(c.stream);
            (void)hipStreamSynchronize(c.copy_stream);
        }
        learned_release();
        DevPool::get().trim();
    } catch (...) {
    }
}

void kq_default_options(KqSolveOptions *opt)
{
    if (!opt)
        return;
    opt->mode = KQ_MODE_MIXED_FP32;
    opt->reliable_delta = 0.0;  // the mode's default
    opt->refine = 1;
    opt->rel_tol = 0.0;
    opt->learned = 0;
    opt->chunk = 0;
    opt->nr = 0;
    opt->nt = 0;
    opt->verbose = 1;
    opt->fork_tol = 0.0;
    opt->learned_tol = 0.0;
}

KuQCDStatus kq_solve_multishift_ex(void **x, void *b, KuQCDInvertParam *param, const KqSolveOptions *opt,
                                   KqSolveReport *rep)
{
    return guarded([&] {
        int slot = -1;
        KuQCDStatus st = check_solve(x, b, param, 1, &slot);
        if (st != KUQCD_SUCCESS)
            return st;
        KqSolveOptions o;
        kq_default_options(&o);
        if (opt)
            o = *opt;
        void *bb[1] = {b};
        return solve_extended(slot, 1, x, bb, *param, o, rep);
    });
}

KuQCDStatus kq_solve_multishift_batched(int nrhs, void **x, void **b, KuQCDInvertParam *param,
                                        const KqSolveOptions *opt, KqSolveReport *rep)
{
    return guarded([&] {
        if (nrhs <= 0) {
            set_error("nrhs must be positive");
            return KUQCD_ERR_INVALID_PARAM;
        }
        if (b == nullptr) {
            set_error("b is nullptr");
            return KUQCD_ERR_NULL_POINTER;
        }
        for (int r = 0; r < nrhs; ++r)
            if (b[r] == nullptr) {
                set_error("b[r] is nullptr");
                return KUQCD_ERR_NULL_POINTER;
            }
        int slot = -1;
        KuQCDStatus st = check_solve(x, b[0], param, nrhs, &slot);
        if (st != KUQCD_SUCCESS)
            return st;
        KqSolveOptions o;
        kq_default_options(&o);
        if (opt)
            o = *opt;
        return solve_extended(slot, nrhs, x, b, *param, o, rep);
    });
}

void *kq_device_alloc(size_t bytes)
{
    try {
        ensure_device();
        void *p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            set_error("hipMalloc failed");
            return nullptr;
        }
        return p;
    } catch (const std::exception &e) {
        set_error(e.what());
        return nullptr;
    }
}
void kq_device_free(void *p)
{
    if (p)
        (void)hipFree(p);
}
int kq_synchronize(void)
{
    try {
        ensure_device();
        KQ_HIP_CHECK(hipStreamSynchronize(ctx().stream));
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}
int kq_memcpy(void *dst, const void *src, size_t bytes)
{
    try {
        ensure_device();
        KQ_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ctx().stream));
        KQ_HIP_CHECK(hipStreamSynchronize(ctx().stream));
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}

const char *kq_last_error(void)
{
    return ctx().last_error.c_str();
}
void kq_set_verbosity(int v)
{
    ctx().verbosity = v;
}
int kq_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess)
        return 0;
    return n;
}

}  // extern "C"

// ================================================================ C++ API (drop-in)
namespace kuqcd {
KuQCDStatus kuqcd_init()
{
    return kq_init();
}
KuQCDStatus kuqcd_load_gaugefield(void *h_gauge, KuQCDGaugeParam *param)
{
    return kq_load_gaugefield(h_gauge, param);
}
KuQCDStatus kuqcd_solve_multishift(void **_hp_x, void *_hp_b, KuQCDInvertParam *param)
{
    return kq_solve_multishift(_hp_x, _hp_b, param);
}
void kuqcd_gauge_free()
{
    kq_gauge_free();
}
void kuqcd_end()
{
    kq_end();
}
}  // namespace kuqcd
