// Library state behind the C and C++ APIs: the HIP stream and the loaded
// gauge fields in their device formats. One host thread drives the library
// (the reference API is likewise not re-entrant, kuqcd.cpp:17-38).
#pragma once

#include <map>
#include <string>
#include <tuple>

#include "kq_solver.h"
#include "kuqcd_enum.h"

namespace kq {

// A loaded link field (fat or long) in the precision it was loaded in,
// pre-scaled (1/2 fat, -1/48 long): the canonical SoA copy
// canon[((q*4 + mu)*9 + k)*Vh + s] from which every kernel format is made,
// and the kernel format of the loaded precision (LinkD or LinkF).
struct GaugeSlot {
    bool loaded = false;
    int prec = 0;  // 4 or 8
    unsigned lat[4] = {0, 0, 0, 0};
    DevBuf soa;  // canonical, double2 or float2
    DevBuf kf;   // LinkD (double) or LinkF (float) format
};

// Lower-precision kernel formats of a double gauge pair, made on first use.
struct DerivedFormats {
    bool f32 = false, f16 = false;
    bool long12 = false;  // long links stored as 12 reals (LinkH12), else LinkH
    DevBuf fat_f, lng_f;  // LinkF
    DevBuf fat_h, lng_h;  // LinkH / LinkH12
};

inline LinkF as_linkF(const DevBuf &b, uint32_t Vh) { return LinkF{b.as<char>(), nb64_of(Vh)}; }
inline LinkH as_linkH(const DevBuf &b, uint32_t Vh) { return LinkH{b.as<char>(), nb64_of(Vh)}; }
inline LinkH12 as_linkH12(const DevBuf &b, uint32_t Vh) { return LinkH12{b.as<char>(), nb64_of(Vh)}; }

struct Context {
    bool device_ready = false;
    hipStream_t stream = nullptr;
    int grid[4] = {1, 1, 1, 1};
    // slot order = reference dispatch order (kuqcd.cpp:415-435):
    // 0 float/BATCH, 1 double/BATCH, 2 float/NORMAL, 3 double/NORMAL
    GaugeSlot fat[4], lng[4];
    DerivedFormats der[4];
    int verbosity = 1;
    std::string last_error;
    // autotuner cache: (mode, nrhs, Vh) -> (nr, nt)
    std::map<std::tuple<int, int, uint32_t>, std::pair<int, int>> tune;
};

Context &ctx();
void ensure_device();
inline int slot_index(int order, int prec) { return (order == KUQCD_BATCH_ORDER ? 0 : 2) + (prec == KUQCD_DOUBLE ? 1 : 0); }
// first slot with both fat and long loaded, in dispatch order; -1 if none
int active_slot();
Geom slot_geom(int slot);

LinkD link_d(const GaugeSlot &s);
LinkF link_f(const GaugeSlot &s);
// derived formats of a double slot (created on demand)
DerivedFormats &derived(int slot, bool want_f32, bool want_f16);

void log_info(const char *fmt, ...);
void set_error(const std::string &msg);

}  // namespace kq
