// Conversions between the reference's host layouts and the device formats.
//
// Host gauge layout (kuqcd.cpp:103-127, NORMAL and BATCH order alike):
//   h[2*(l + 9*(mu + 4*i)) + ri], i = q*Vh + s over the whole lattice, l = 3*row + col.
// Host colour-vector layout (kuqcd.cpp:129-169):
//   h[2*(c + 3*s) + ri], s over the Vh sites of the solve parity.
#pragma once

#include "kq_fields.h"

namespace kq {

// raw: device copy of host-layout links for full-lattice sites [i0, i0 + n).
// dst: SoA complex links (9 per link) for the whole lattice; every entry is
// multiplied by `scale` (1/2 for fat links, -1/48 for long links).
template <typename HR, typename LR>
void gauge_host_to_soa(const HR *raw, size_t i0, size_t n, c2_t<LR> *dst, uint32_t Vh, double scale,
                       hipStream_t st);

// Kernel link formats (kq_fields.h) from the canonical SoA links
// canon[((q*4 + mu)*9 + k)*Vh + s] (complex):
// FP64 -> LinkD; FP64 or FP32 -> LinkF (rounding to nearest)
void canon_to_linkD(const double2 *canon, char *dst, uint32_t Vh, hipStream_t st);
// FP64 long links as LinkD14; *fail (device int, not cleared here) counts the links that are not
// a complex multiple of an SU(3) matrix to 1e-13
void canon_to_linkD14(const double2 *canon, char *dst, uint32_t Vh, int *fail, hipStream_t st);
template <typename CR> void canon_to_linkF(const c2_t<CR> *canon, char *dst, uint32_t Vh, hipStream_t st);
// FP64 -> LinkH: 16-bit fixed-point entries with one FP32 scale per link (max |entry| / 32767)
void canon_to_linkH(const double2 *canon, char *dst, uint32_t Vh, hipStream_t st);
// FP64 links of the form b * V (V in SU(3), b real) -> LinkH12: rows 0,1 of V in
// 16-bit fixed point and b in FP32. fail[0] counts links that are not of that form to 1e-6.
void canon_to_linkH12(const double2 *canon, char *dst, uint32_t Vh, int *fail, hipStream_t st);
// LinkF12 from canonical double links; *fail counts links that are not b * SU(3)
void canon_to_linkF12(const double2 *canon, char *dst, uint32_t Vh, int *fail, hipStream_t st);
void canon_to_linkF12(const float2 *canon, char *dst, uint32_t Vh, int *fail, hipStream_t st);
// long links as complex multiples of SU(3) matrices (LinkF12C, LinkH12C); *fail counts the links
// without that form to 1e-6
void canon_to_linkF12C(const double2 *canon, char *dst, uint32_t Vh, int *fail, hipStream_t st);
void canon_to_linkF12C(const float2 *canon, char *dst, uint32_t Vh, int *fail, hipStream_t st);
void canon_to_linkH12C(const double2 *canon, char *dst, uint32_t Vh, int *fail, hipStream_t st);

// host-layout colour vector (device copy) <-> SoA vector of right-hand side r
template <typename HR, typename VR>
void vec_host_to_soa(const HR *raw, Vec<VR> v, int r, hipStream_t st);
template <typename VR, typename HR>
void vec_soa_to_host(Vec<VR> v, int r, HR *raw, hipStream_t st);

// y = x with precision conversion (same SoA layout, n complex numbers)
template <typename XR, typename YR>
void convert_copy(const c2_t<XR> *x, c2_t<YR> *y, size_t n, hipStream_t st);

// plain device-to-device copy of `bytes` (multiple of 16) with 16-byte loads and
// stores: the HBM reference point for the roofline
void device_copy(const void *src, void *dst, size_t bytes, hipStream_t st);

}  // namespace kq
