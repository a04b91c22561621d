// Learned-preconditioner refinement used by the mixed-precision solve.
#pragma once

#include "kq_context.h"
#include "kuqcd_hip.h"

namespace kq {

struct RefineResult {
    int iter = 0, iter_learned = 0, reliable = 0;
    bool fallback = false, converged = false;
    double rel0 = 0.0, rel = 0.0;  // FP64 true relative residual before and after
    double seconds = 0.0;
};

// true once weights were set or trained
bool learned_weights_available();

// Flexible-CG refinement of x (device SoA, double2[3*Vh]) for A = sigma - D D on
// `parity` with the gauge of `slot`; b likewise. Options as kq_fcg_refine.
RefineResult fcg_refine_device(int slot, int parity, double sigma, const double *b, double *x,
                               const KqRefineOptions &o);

// free the training set (device memory)
void learned_release();

}  // namespace kq
