// Solve drivers behind kq_solve_multishift / _ex / _batched.
#pragma once

#include "kq_context.h"
#include "kuqcd_hip.h"

namespace kq {

// The reference method (CGM::Invert) in the loaded gauge precision.
KuQCDStatus solve_reference(int slot, void **x, void *b, KuQCDInvertParam &p);

// FP64 or mixed-precision multishift CG for nrhs sources with per-shift FP64
// verification and refinement.
KuQCDStatus solve_extended(int slot, int nrhs, void **x, void **b, KuQCDInvertParam &p, const KqSolveOptions &o,
                           KqSolveReport *rep);

// sum_blk part[k][blk] on the host side
std::vector<double> reduce_host(const double *part_dev, uint32_t nblk, int n, hipStream_t st);

// autotuned (nr, nt) for the hopping term
std::pair<int, int> tuned_launch(int slot, int mode, int nrhs);

}  // namespace kq
