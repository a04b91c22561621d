// Learned preconditioner: a two-layer parallel-transport convolution on the
// even-odd lattice, built from the same fat and long links as the operator.
//
//   h_c(y) = sum_t W1[c][t] (T_t r)(y)                      y on the other parity, c < 4
//   z(x)   = W0 r(x) + sum_t T_t( sum_c W2[c][t] h_c )(x)   x on the parity of r
//
// with 16 taps t = (mu, sign, hop in {1, 3}):
//   (T_{mu,+,h} v)(x) = L_h,mu(x) v(x + h mu),   (T_{mu,-,h} v)(x) = L_h,mu(x - h mu)^+ v(x - h mu)
// (L_1 = fat link, L_3 = long link, as stored for the operator, i.e. with its
// 1/2 and -1/48 factors). Every tap maps a covariant field to a covariant
// field, so the network commutes with gauge transformations. The adjoint of
// tap t is the tap with the opposite sign, which gives the backward pass.
//
// The trainable weights are complex FP32 master values in units of the
// gauge field's link norms: the kernels use w_eff = bf16(w / nu_h), where
// nu_h is the root-mean-square Frobenius norm per colour row of the hop-h
// links, so that every tap has unit scale whatever the normalisation of the
// links. Kernels read w_eff as packed bfloat16 pairs and accumulate in FP32.
#pragma once

#include "kq_dslash.h"

namespace kq {

constexpr int kPtTaps = 16;
constexpr int kPtHidden = 4;
// parameter count in complex numbers: W1 [4][16], W2 [4][16], W0
constexpr int kPtParams = 2 * kPtHidden * kPtTaps + 1;

inline int pt_tap_index(int mu, bool negative, int hop) { return mu * 4 + (negative ? 2 : 0) + (hop == 3 ? 1 : 0); }
inline int pt_tap_hop(int t) { return (t & 1) ? 3 : 1; }
inline int pt_flip_tap(int t) { return t ^ 2; }  // opposite sign, same direction and hop

// tap t = mu*4 + (negative ? 2 : 0) + (hop 3 ? 1 : 0)
struct PtWeights {
    float2 w1[kPtHidden][kPtTaps];
    float2 w2[kPtHidden][kPtTaps];
    float2 w0;
};

// Weights as the kernels take them: bfloat16 real part in the low half,
// imaginary part in the high half of each word.
struct PtWeightsBF {
    uint32_t w1[kPtHidden][kPtTaps];
    uint32_t w2[kPtHidden][kPtTaps];
    uint32_t w0;
};

// Link-norm scales of a gauge pair: nu[0] for the fat (hop 1), nu[1] for the long (hop 3) links.
struct PtNorms {
    double nu[2] = {1.0, 1.0};
};

float bf16_round(float f);
uint32_t bf16_pack(float2 w);
float2 bf16_unpack(uint32_t u);
// w_eff = bf16(w / nu_h), packed
PtWeightsBF pt_pack(const PtWeights &w, const PtNorms &nu);
// packed weights of the hidden-layer backward pass: w1'[c][t] = conj(w2_eff[c][flip t])
PtWeightsBF pt_pack_backward(const PtWeights &w, const PtNorms &nu);
// effective weights unpacked to FP32 (what the kernels compute with)
PtWeights pt_effective(const PtWeights &w, const PtNorms &nu);

// Forward: h (4 channels, channel stride 3*Vh, parity 1-q) and z (parity q) from r (parity q).
// If dotq is given, also writes per-block partial sums of <dotq, z> (complex) to
// part[2*blk], part[2*blk+1], with pt_layer_nblk(Vh) blocks. If stop is given, the
// kernels do nothing while *stop != 0.
// If h16 is given, the hidden channels pass from layer 1 to layer 2 as bfloat16 in h16
// (3*Vh entries of 16 B: colour k of site s holds the 4 channels' (re, im) bfloat16 pairs
// at h16[k*Vh + s]); h (FP32) is then written only if not null. Otherwise layer 2 reads h.
template <class LF, class LL>
void pt_forward(const Geom &g, int q, LF fat, LL lng, const PtWeightsBF &w, const float2 *r, float2 *h, float2 *z,
                hipStream_t st, const float2 *dotq = nullptr, double *part = nullptr, const int *stop = nullptr,
                kq_v4u *h16 = nullptr);
// FP32 copy of bfloat16 hidden channels (h16 layout above -> 4 channels, channel stride 3*Vh)
void pt_unpack_hidden(const kq_v4u *h16, float2 *h, uint32_t Vh, hipStream_t st);
// Backward to the hidden layer: gh_c = sum_t conj(W2[c][t]) T_t^+ g  (parity 1-q); wb from pt_pack_backward
template <class LF, class LL>
void pt_backward_hidden(const Geom &g, int q, LF fat, LL lng, const PtWeightsBF &wb, const float2 *gz, float2 *gh,
                        hipStream_t st);
// Gradients with respect to the conjugate effective weights for one sample, in FP64:
//   gW1[c][t] = sum_y <T_t r, gh_c>,  gW2[c][t] = sum_y <h_c, T_t^+ g>,  gW0 = sum_x <r, g>
// out[2*j], out[2*j+1] = re, im of parameter j (order w1, w2, w0)
template <class LF, class LL>
void pt_weight_grads(const Geom &g, int q, LF fat, LL lng, const float2 *r, const float2 *h, const float2 *gz,
                     const float2 *gh, double *out, hipStream_t st);
uint32_t pt_layer_nblk(uint32_t Vh);

// ---- FP32 vector helpers (n complex entries)
void vf_axpy(float2 *y, const float2 *x, float2 a, size_t n, hipStream_t st);  // y += a x (complex a)
void vf_lincomb(float2 *y, const float2 *x1, float2 a1, const float2 *x2, float2 a2, size_t n,
                hipStream_t st);                                                   // y = a1 x1 + a2 x2
// several complex dot products <x_i, y_i> (sum over n complex entries) in one launch; out: host
void vf_dots(const float2 *const *xs, const float2 *const *ys, int ndot, size_t n, double2 *out, hipStream_t st);
// sum of squares of n doubles (host result)
double vd_sumsq(const double *x, size_t n, hipStream_t st);

}  // namespace kq
