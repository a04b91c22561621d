// Library state behind the C and C++ APIs: the HIP stream and the loaded
// gauge fields in their device formats. One host thread drives the library
// (the reference API is likewise not re-entrant, kuqcd.cpp:17-38).
#pragma once

#include <array>
#include <map>
#include <string>
#include <tuple>

#include "kq_solver.h"
#include "kuqcd_enum.h"
#include "kuqcd_hip.h"

namespace kq {

// A loaded link field (fat or long) in the precision it was loaded in,
// pre-scaled (1/2 fat, -1/48 long): the canonical SoA copy
// canon[((q*4 + mu)*9 + k)*Vh + s] from which every kernel format is made,
// and the kernel format of the loaded precision (LinkD or LinkF).
// Long links in fewer reals: a real multiple of an SU(3) matrix (kLong12), a complex one
// (kLong12C, U(3) products), else 18 reals (kLong18)
enum LongForm { kLong18 = 0, kLong12 = 1, kLong12C = 2 };

struct GaugeSlot {
    bool loaded = false;
    int prec = 0;  // 4 or 8
    unsigned lat[4] = {0, 0, 0, 0};
    DevBuf soa;  // canonical, double2 or float2
    DevBuf kf;   // LinkD (double) or LinkF (float) format
    // long links also held in fewer reals in kfr when every link has that form: kLong12C in
    // double (LinkD14, a complex multiple of an SU(3) matrix); kLong12 (LinkF12, a real multiple)
    // or kLong12C (LinkF12C) in float
    int lform = kLong18;
    DevBuf kfr;
};

// Lower-precision kernel formats of a double gauge pair, made on first use.
struct DerivedFormats {
    bool f32 = false, f16 = false;
    int h_long = kLong18;  // 16-bit long links in lng_h: LinkH, LinkH12 or LinkH12C
    int f_long = kLong18;  // FP32 long links: LinkF (lng_f), or LinkF12 / LinkF12C (lng_f12)
    DevBuf fat_f, lng_f;   // LinkF
    DevBuf lng_f12;        // LinkF12 or LinkF12C
    DevBuf fat_h, lng_h;   // LinkH / LinkH12 / LinkH12C
};

inline LinkF as_linkF(const DevBuf &b, uint32_t Vh) { return LinkF{b.as<char>(), nb64_of(Vh)}; }
inline LinkH as_linkH(const DevBuf &b, uint32_t Vh) { return LinkH{b.as<char>(), nb64_of(Vh)}; }
inline LinkH12 as_linkH12(const DevBuf &b, uint32_t Vh) { return LinkH12{b.as<char>(), nb64_of(Vh)}; }
inline LinkF12 as_linkF12(const DevBuf &b, uint32_t Vh) { return LinkF12{b.as<char>(), nb64_of(Vh)}; }
inline LinkF12C as_linkF12C(const DevBuf &b, uint32_t Vh) { return LinkF12C{b.as<char>(), nb64_of(Vh)}; }
inline LinkH12C as_linkH12C(const DevBuf &b, uint32_t Vh) { return LinkH12C{b.as<char>(), nb64_of(Vh)}; }

struct Context {
    bool device_ready = false;
    hipStream_t stream = nullptr;
    // second stream (default priority) for device-to-host copies that overlap the solver's kernels
    hipStream_t copy_stream = nullptr;
    int grid[4] = {1, 1, 1, 1};
    // slot order = reference dispatch order (kuqcd.cpp:415-435):
    // 0 float/BATCH, 1 double/BATCH, 2 float/NORMAL, 3 double/NORMAL
    GaugeSlot fat[4], lng[4];
    DerivedFormats der[4];
    int verbosity = 1;
    std::string last_error;
    // autotuner cache: (mode, nrhs, Vh) -> (nr, nt)
    std::map<std::tuple<int, int, uint32_t>, std::pair<int, int>> tune;
    // (mode, Vh) -> milliseconds of one group of NR = 1, 2, 3, 4, 8 right-hand sides at block
    // size 64, 128, 256 (index 3 * (position of NR in that list) + log2(nt / 64)); < 0 where
    // not measured
    std::map<std::pair<int, uint32_t>, std::array<double, 15>> tune_group;
    // (mode, Vh, nr, nt) -> occupancy cap of the hopping term (waves per SIMD, 0 = none)
    std::map<std::tuple<int, uint32_t, int, int>, int> tune_cap;
};

Context &ctx();
void ensure_device();
inline int slot_index(int order, int prec) { return (order == KUQCD_BATCH_ORDER ? 0 : 2) + (prec == KUQCD_DOUBLE ? 1 : 0); }
// first slot with both fat and long loaded, in dispatch order; -1 if none
int active_slot();
Geom slot_geom(int slot);

LinkD link_d(const GaugeSlot &s);
LinkD14 link_d14(const GaugeSlot &s);
// fn(LinkD fat, lng) with the long links of the slot's double pair as LinkD14 when every one
// has that form (GaugeSlot::lform), else as LinkD: the FP64 hopping term of the solvers, the FP64
// verification and the operator entry points
template <class Fn> inline void with_fp64_links(int slot, Fn &&fn);
LinkF link_f(const GaugeSlot &s);
// derived formats of a double slot (created on demand)
DerivedFormats &derived(int slot, bool want_f32, bool want_f16);

// fn(fat, lng) with the device link formats of a mixed mode of a double slot:
// KQ_MODE_MIXED_FP32 -> LinkF fat; long links LinkF12 (b * SU(3)), LinkF12C (w * SU(3), w complex)
//                       or LinkF;
// KQ_MODE_MIXED_FP16 -> LinkH fat; long links LinkH12, LinkH12C or LinkH;
// KQ_MODE_MIXED_FP32_L16 -> LinkF fat, long links as in KQ_MODE_MIXED_FP16.
inline bool is_mixed_mode(int mode)
{
    return mode == KQ_MODE_MIXED_FP32 || mode == KQ_MODE_MIXED_FP16 || mode == KQ_MODE_MIXED_FP32_L16;
}

template <class FAT, class Fn> void with_h_long(const DerivedFormats &d, FAT fat, uint32_t Vh, Fn &&fn)
{
    if (d.h_long == kLong12)
        fn(fat, as_linkH12(d.lng_h, Vh));
    else if (d.h_long == kLong12C)
        fn(fat, as_linkH12C(d.lng_h, Vh));
    else
        fn(fat, as_linkH(d.lng_h, Vh));
}

template <class Fn> void with_mixed_links(int slot, int mode, Fn &&fn)
{
    const uint32_t Vh = slot_geom(slot).Vh;
    if (mode == KQ_MODE_MIXED_FP32_L16) {
        DerivedFormats &d = derived(slot, true, true);
        with_h_long(d, as_linkF(d.fat_f, Vh), Vh, fn);
        return;
    }
    if (mode == KQ_MODE_MIXED_FP16) {
        DerivedFormats &d = derived(slot, false, true);
        with_h_long(d, as_linkH(d.fat_h, Vh), Vh, fn);
        return;
    }
    if (mode != KQ_MODE_MIXED_FP32)
        throw std::invalid_argument("link format must be a mixed mode (KQ_MODE_MIXED_*)");
    DerivedFormats &d = derived(slot, true, false);
    if (d.f_long == kLong12)
        fn(as_linkF(d.fat_f, Vh), as_linkF12(d.lng_f12, Vh));
    else if (d.f_long == kLong12C)
        fn(as_linkF(d.fat_f, Vh), as_linkF12C(d.lng_f12, Vh));
    else
        fn(as_linkF(d.fat_f, Vh), as_linkF(d.lng_f, Vh));
}

template <class Fn> inline void with_fp64_links(int slot, Fn &&fn)
{
    Context &c = ctx();
    if (c.lng[slot].lform != kLong18)
        fn(link_d(c.fat[slot]), link_d14(c.lng[slot]));
    else
        fn(link_d(c.fat[slot]), link_d(c.lng[slot]));
}

// fn(LinkF fat, lng) for a float pair: long links as LinkF12 or LinkF12C when every one is a real
// or a complex multiple of an SU(3) matrix (GaugeSlot::lform), else as LinkF (the float reference
// method and operators)
template <class Fn> inline void with_fp32_links(int slot, Fn &&fn)
{
    Context &c = ctx();
    const GaugeSlot &l = c.lng[slot];
    const uint32_t Vh = (uint32_t)((uint64_t)l.lat[0] * l.lat[1] * l.lat[2] * l.lat[3] / 2);
    if (l.lform == kLong12)
        fn(link_f(c.fat[slot]), as_linkF12(l.kfr, Vh));
    else if (l.lform == kLong12C)
        fn(link_f(c.fat[slot]), as_linkF12C(l.kfr, Vh));
    else
        fn(link_f(c.fat[slot]), link_f(l));
}

// device address of the no-allocation entry point's static scratch array on the current
// device (looked up once per device)
void *noalloc_scratch();

void log_info(const char *fmt, ...);
void set_error(const std::string &msg);

}  // namespace kq
