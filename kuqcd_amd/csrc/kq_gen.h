// Seeded synthetic gauge fields and sources in the reference host layouts.
#pragma once
#include <cstdint>

namespace kq {

enum { KQ_GEN_HAAR = 0, KQ_GEN_WEAK = 1 };

template <typename HR>
int gen_gauge(HR *fat, HR *lng, const unsigned lat[4], int kind, double eps, double fat_scale, double c_naik,
              uint64_t seed_fat, uint64_t seed_long, int phases);
template <typename HR> int gen_source(HR *b, const unsigned lat[4], int parity, uint64_t seed);

}  // namespace kq
