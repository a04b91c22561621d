/*
 * C ABI of the MI355X KuQCD library (libkuqcd.so): plain pointers and sizes,
 * callable from C, ctypes, cgo, JNI.
 *
 * Part 1 re-exports the reference API (include/kuqcd.h:29-34) with C linkage.
 * Part 2 adds the MI355X extensions: precision modes, per-shift FP64
 * verification and refinement, multi-right-hand-side batching, operator-level
 * entry points, synthetic inputs and timing hooks. Host buffers always use the
 * reference's host layouts (kuqcd.cpp:103-169):
 *   gauge  : h[2*(l + 9*(mu + 4*i)) + ri], i = parity*Vh + checkerboard index
 *   vector : h[2*(c + 3*s) + ri], s = checkerboard index of one parity
 */
#ifndef KUQCD_HIP_H
#define KUQCD_HIP_H

#include <stddef.h>
#include <stdint.h>
#include "kuqcd_enum.h"

#define KQ_API __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------ Part 1: reference API, C linkage */
/* replaces kuqcd::kuqcd_init (kuqcd.h:29, kuqcd.cpp:288-301): grid from stdin */
KQ_API KuQCDStatus kq_init(void);
/* kuqcd_init with the grid given directly instead of read from stdin */
KQ_API KuQCDStatus kq_init_grid(const int grid[4]);
/* replaces kuqcd::kuqcd_load_gaugefield (kuqcd.h:30, kuqcd.cpp:303-402) */
KQ_API KuQCDStatus kq_load_gaugefield(void *h_gauge, KuQCDGaugeParam *param);
/* replaces kuqcd::kuqcd_solve_multishift (kuqcd.h:31, kuqcd.cpp:404-445) */
KQ_API KuQCDStatus kq_solve_multishift(void **x, void *b, KuQCDInvertParam *param);
/* replaces kuqcd::kuqcd_gauge_free (kuqcd.h:33, kuqcd.cpp:447-476) */
KQ_API void kq_gauge_free(void);
/* replaces kuqcd::kuqcd_end (kuqcd.h:34, kuqcd.cpp:477-487) */
KQ_API void kq_end(void);

/* ------------------------------------------------ Part 2: extensions (no reference counterpart) */
typedef enum {
    KQ_MODE_REFERENCE = 0,  /* the reference's method in the loaded gauge precision */
    KQ_MODE_FP64 = 1,       /* plain FP64 multishift CG (double gauge) */
    KQ_MODE_MIXED_FP32 = 2, /* FP32 links/vectors, FP64 reliable updates + per-shift refinement */
    KQ_MODE_MIXED_FP16 = 3, /* 16-bit links: fixed-point int16 entries with a per-link FP32 scale
                               (12-real long links), FP32 vectors */
    KQ_MODE_MIXED_FP32_L16 = 4 /* FP32 fat links, 16-bit long links (the Naik term is 1/24 of the
                                  fat term, so its rounding costs the solve little), FP32 vectors */
} KqMode;

typedef struct {
    int mode;              /* KqMode */
    double reliable_delta; /* reliable update when |r| drops below delta * |r| at the last one;
                              0 (default): 0.02, or 0.1 with 16-bit fat links (KQ_MODE_MIXED_FP16) */
    int refine;            /* 1: FP64 true-residual check of every shift + batched refinement */
    double rel_tol;        /* per-shift target of ||b - A_k x_k|| / ||b||; 0: sqrt(tol)/||b|| */
    int learned;           /* 1: learned-preconditioner flexible CG for the lightest shift */
    int chunk;             /* iterations replayed between host checks (0: default) */
    int nr;                /* right-hand sides per thread of the batched hopping term (0: autotune) */
    int nt;                /* threads per block of the hopping term (0: autotune) */
    int verbose;           /* 0 silent, 1 iteration summary (reference print), 2 details */
    double fork_tol;       /* mixed modes: a shift k >= 1 leaves the multishift and continues as its own
                              FP64-reliable CG once its iterated relative residual is below fork_tol
                              (0: automatic, 1e-7; < 0: never) */
    double learned_tol;    /* learned = 1: relative residual at which shift 0 leaves the multishift for
                              the flexible-CG refinement (0: 1e-6) */
} KqSolveOptions;

#define KQ_MAX_SHIFT 32
typedef struct {
    int iter;                          /* multishift iterations (reference's param->iter) */
    int reliable_updates;
    int refine_rounds;
    int refine_iter[KQ_MAX_SHIFT];     /* refinement iterations per shift */
    double true_rel_res[KQ_MAX_SHIFT]; /* FP64 ||b - A_k x_k|| / ||b|| per shift */
    double seconds;                    /* wall time of the whole call, host buffers in and out */
    double seconds_multishift;
    double seconds_refine;
    int learned_used;                  /* 1 if the learned refinement was accepted */
    int learned_fallback;              /* 1 if it was rejected and plain CG refinement used */
    double initial_rel_res[KQ_MAX_SHIFT]; /* FP64 true relative residual after the multishift phase */
    double seconds_verify;             /* FP64 true-residual checks */
} KqSolveReport;

KQ_API void kq_default_options(KqSolveOptions *opt);

/* Multishift solve with options; same inputs and shift/tol conventions as
 * kq_solve_multishift. rep may be NULL. */
KQ_API KuQCDStatus kq_solve_multishift_ex(void **x, void *b, KuQCDInvertParam *param, const KqSolveOptions *opt,
                                   KqSolveReport *rep);

/* nrhs independent sources solved together (one gauge field): b[r] host
 * vectors, x[r*num_shift + k] host solutions, param->shifts holds
 * nrhs*num_shift shifts (reference convention per source). rep: nrhs reports (may be NULL). */
KQ_API KuQCDStatus kq_solve_multishift_batched(int nrhs, void **x, void **b, KuQCDInvertParam *param,
                                        const KqSolveOptions *opt, KqSolveReport *rep);

/* dst = D src (hopping term, dirac_hisq.cpp:409-455) with the loaded gauge:
 * src on parity 1-dst_parity, host buffers in the gauge precision;
 * mode picks the device link/vector format (KqMode; REFERENCE = as loaded). */
KQ_API KuQCDStatus kq_dslash(const void *src, void *dst, int dst_parity, int mode);
/* nrhs vectors back to back (host layout each); nr/nt as in KqSolveOptions */
KQ_API KuQCDStatus kq_dslash_batched(int nrhs, const void *src, void *dst, int dst_parity, int mode, int nr, int nt);
/* dst = sigma*src - D D src on parity `parity` (the CGM operator A) */
KQ_API KuQCDStatus kq_mdaggm(const void *src, void *dst, int parity, double sigma, int mode);
/* No counterpart in the reference. For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): kq_mdaggm in
 * KQ_MODE_REFERENCE (the loaded gauge's precision) without any allocation. src and dst are
 * device memory in the reference host layout (Vh*6 reals each, not overlapping); dst also holds
 * the intermediate D src during the call. The library's scratch for it is one static device
 * array of KQ_NOALLOC_SCRATCH_BYTES per GPU (the FP64 operator up to 32^4 sites); a larger
 * lattice returns KUQCD_ERR_INVALID_PARAM. The work is queued on the library stream and not
 * waited for (kq_synchronize waits). Calls that overlap on one GPU race for the scratch. */
#define KQ_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
KQ_API KuQCDStatus kq_mdaggm_noalloc(const void *src, void *dst, int parity, double sigma);

/* ---- timing hooks (device-resident inputs, HIP events on the library stream) */
typedef struct {
    double ms;             /* average time per application */
    double bytes;          /* algorithmic bytes per application */
    double gbs;            /* bytes / ms */
    double gflops;         /* 1146 flop per site per right-hand side */
    double hbm_frac;       /* gbs / 8000 */
    int nr, nt;            /* launch configuration used */
} KqBenchResult;
KQ_API KuQCDStatus kq_bench_dslash(int mode, int nrhs, int nr, int nt, int iters, KqBenchResult *out);
/* plain device copy of `bytes` bytes (HBM reference point) */
KQ_API KuQCDStatus kq_bench_copy(size_t bytes, int iters, KqBenchResult *out);
/* pick nr, nt for (mode, nrhs) on the loaded lattice by timing; cached */
KQ_API KuQCDStatus kq_autotune_dslash(int mode, int nrhs, int *nr, int *nt);

/* ---- synthetic inputs (reference host layouts, double precision) */
/* kind 0: Haar-random SU(3); kind 1: exp(i eps H), H traceless Hermitian Gaussian.
 * fat = fat_scale * eta * G(seed_fat); long = c_naik * eta * W(x)W(x+mu)W(x+2mu), W = G(seed_long);
 * either output may be NULL. phases = 1 folds the staggered phases in. */
KQ_API int kq_gen_gauge(double *fat, double *lng, const unsigned lat[4], int kind, double eps, double fat_scale,
                 double c_naik, uint64_t seed_fat, uint64_t seed_long, int phases);
/* complex Gaussian source, E|z|^2 = 1, on the Vh sites of `parity` */
KQ_API int kq_gen_source(double *b, const unsigned lat[4], int parity, uint64_t seed);

/* ---- device memory for callers that keep b and x on the GPU: every entry
 * point above accepts host or device pointers (reference layout either way) */
KQ_API void *kq_device_alloc(size_t bytes);
KQ_API void kq_device_free(void *p);
/* copy `bytes` between any two host/device pointers; returns 0 on success */
KQ_API int kq_memcpy(void *dst, const void *src, size_t bytes);
/* wait until all work the library has queued on its stream is done; returns 0 on success */
KQ_API int kq_synchronize(void);

/* ---- learned preconditioner (no reference counterpart)
 * A two-layer parallel-transport convolution M built from the loaded fat and
 * long links: 16 taps (4 directions x forward/backward x hops 1 and 3), 4
 * hidden colour-vector channels, complex weights applied in bfloat16 with
 * FP32 accumulation:
 *   h_c = sum_t W1[c][t] T_t r  (other parity),  z = W0 r + sum_t T_t sum_c W2[c][t] h_c
 * tap t = mu*4 + 2*(backward) + (hop 3). Weights are in units of the links'
 * root-mean-square row norm nu (per hop), i.e. the kernels use bf16(W / nu_hop).
 * Trained to minimise the error of the correction z = M r in the A^2 norm,
 *   L(r, e) = |A (z - e)|^2 / |A e|^2,  e = A^{-1} r  (= |r - A M r|^2 / |r|^2 for exact pairs),
 * on (residual, error) pairs that the solver generates. */
typedef struct {
    float w1[4][16][2]; /* [channel][tap][re, im] */
    float w2[4][16][2];
    float w0[2];
} KqPtWeights;

/* W0 = 1, W1 and W2 complex Gaussian with standard deviation `scale` (host only) */
KQ_API void kq_ptnet_init_weights(KqPtWeights *w, uint64_t seed, double scale);
/* weights used by the learned refinement (kq_solve_multishift_ex with learned = 1, kq_fcg_refine) */
KQ_API KuQCDStatus kq_ptnet_set_weights(const KqPtWeights *w);
KQ_API KuQCDStatus kq_ptnet_get_weights(KqPtWeights *w);
/* link norms nu[0] (fat), nu[1] (long) of the loaded double gauge pair */
KQ_API KuQCDStatus kq_ptnet_norms(double nu[2]);
/* z = M r on parity `parity` with the loaded gauge (r, z: reference layout, double, host or device);
 * h (may be NULL): the 4 hidden channels, each a reference-layout vector on parity 1-parity.
 * link_format: KQ_MODE_MIXED_FP32 or KQ_MODE_MIXED_FP16 */
KQ_API KuQCDStatus kq_ptnet_forward(const KqPtWeights *w, int parity, int link_format, const void *r, void *z,
                                    void *h);
/* the same with the hidden channels passed from layer 1 to layer 2 as bfloat16 (round to nearest
 * even), as the flexible CG applies the network; h (may be NULL) receives the rounded channels */
KQ_API KuQCDStatus kq_ptnet_forward_h16(const KqPtWeights *w, int parity, int link_format, const void *r, void *z,
                                        void *h);
/* loss L(r, e) for A = sigma - D D and its gradient dL/dconj(W) with respect to the weights as
 * stored in KqPtWeights (258 doubles: w1, w2, w0 as re, im pairs); FP32 links and vectors */
KQ_API KuQCDStatus kq_ptnet_loss_grad(const KqPtWeights *w, int parity, double sigma, const void *r, const void *e,
                                      double *loss, double *grad);

/* Training set: each call adds the loaded gauge pair (FP32 copy) and, for `nsrc` Gaussian sources
 * (seeds seed, seed+1, ...), the pairs (r_k, x* - x_k) of an FP64-reliable CG on A = sigma - D D
 * from x_0 = 0 down to relative residual rel_tol, taken at every reliable update whose relative
 * residual is at least min_level; vectors are normalised to |r_k| = 1. *nsamples: total so far. */
KQ_API KuQCDStatus kq_learned_collect(int parity, double sigma, int nsrc, uint64_t seed, double rel_tol,
                                      double min_level, int *nsamples);
KQ_API void kq_learned_clear(void);
typedef struct {
    int steps;        /* optimizer steps */
    int batch;        /* samples per step */
    double lr;        /* Adam step size at the start (cosine decay to lr_final) */
    double lr_final;
    uint64_t seed;    /* sample order and initial noise */
    int init;         /* 0: start from the current weights; 1: kq_ptnet_init_weights(seed, init_scale);
                         2: channel 0 = the degree-1 polynomial 8/b - 0.9*8/b^2 A of A (b = 1.02 x its
                         largest eigenvalue by power iteration on the first training gauge), the other
                         channels random with standard deviation init_scale (0: 0.01) */
    double init_scale;
    int verbose;
} KqTrainOptions;
#define KQ_TRAIN_HIST 64
typedef struct {
    int steps, nsamples, ngauges;
    double loss_initial;            /* mean loss over the training set before training */
    double loss_final;              /* ... and after */
    double loss_hist[KQ_TRAIN_HIST]; /* mean batch loss, sampled evenly over the steps */
    double seconds;
} KqTrainReport;
/* Adam on the training set; the result becomes the current weights */
KQ_API KuQCDStatus kq_learned_train(const KqTrainOptions *o, KqTrainReport *rep);
/* mean loss of weights w over the training set */
KQ_API KuQCDStatus kq_learned_eval(const KqPtWeights *w, double *mean_loss);

/* Flexible CG for one shifted system A = sigma - D D on parity `parity` with the loaded gauge:
 * x (reference layout, double, host or device) is refined in place until ||b - A x|| <= rel_tol ||b||
 * (FP64 true residual). FP32 or FP16 links and FP32 vectors inside, FP64 reliable updates every
 * factor `delta` of residual reduction. With learned = 1 the current weights precondition it; if a
 * reliable update finds the FP64 residual not reduced, or no reduction by delta within stall_iter
 * iterations, it falls back to plain CG from the best iterate. */
typedef struct {
    double rel_tol;
    int learned;
    int mode;         /* KQ_MODE_MIXED_FP32 or KQ_MODE_MIXED_FP16: link format of A */
    int net_format;   /* link format of the network (0: FP16) */
    int max_iter;
    double delta;     /* 0: 0.1 */
    int chunk;        /* iterations per graph replay (0: 8) */
    int stall_iter;   /* 0: 300 */
    int verbose;
} KqRefineOptions;
typedef struct {
    int iter;             /* flexible-CG iterations in total */
    int iter_learned;     /* of which with the network */
    int reliable_updates;
    int fallback;         /* 1 if the network was dropped for plain CG */
    int converged;
    double rel_res_initial;
    double true_rel_res;  /* FP64 ||b - A x|| / ||b|| at the end */
    double seconds;       /* device-resident part (without host copies) */
} KqRefineReport;
KQ_API void kq_default_refine_options(KqRefineOptions *o);
KQ_API KuQCDStatus kq_fcg_refine(int parity, double sigma, const void *b, void *x, const KqRefineOptions *o,
                                 KqRefineReport *rep);

/* ---- misc */
KQ_API const char *kq_last_error(void);
KQ_API void kq_set_verbosity(int v);
KQ_API int kq_device_count(void);
/* host check of the fast division used by the kernels' index math: returns
 * the number of n in [0, n_max) with n / d != the fast quotient */
KQ_API long kq_selftest_fastdiv(uint32_t d, uint32_t n_max);
/* host check of the reliable-update fold rule of the mixed solver: 1 if the fold adds the
 * low-precision update of shift k to its FP64 solution, for a system with max_term shifts still
 * iterating and the host's divergence flag `discard` */
KQ_API int kq_selftest_fold_keeps(int discard, int k, int max_term);
/* host evaluation of the mixed solver's segment-fold coefficients (deferred solution updates of
 * a multishift system): h holds J <= 32 recorded iterations, h[(j*4 + f)*32 + k] with f = 0 B_k,
 * 1 Z_k, 2 a_k, and h[(j*4 + 3)*32] = the number of shifts updated in iteration j; out gets 66
 * doubles: [c_p, c_r[0..32), d_p, d_r[0..32)] such that over the segment
 * x_k += c_p p_k(start) + sum_j c_r[j] r_j and p_k(end) = d_p p_k(start) + sum_j d_r[j] r_j */
KQ_API void kq_selftest_seg_coef(const double *h, int J, int k, int keep_x, double *out);
/* checkerboard index of the neighbour of site s (parity q) at hop h along mu,
 * computed by the kernels' index math on the host */
KQ_API long kq_neighbour_index(const int lat[4], int q, long s, int mu, int h);

#ifdef __cplusplus
}
#endif

#endif /* KUQCD_HIP_H */
