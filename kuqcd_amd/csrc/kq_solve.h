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

// bytes from device memory d to h (pageable host, pinned host or device memory), synchronous with
// respect to the host; pageable destinations go through pinned chunks (kq_xfer.cpp)
void copy_to_host(void *h, const void *d, size_t bytes, hipStream_t st);
// bytes from h (pageable host, pinned host or device memory) to device memory d, likewise
void copy_from_host(void *d, const void *h, size_t bytes, hipStream_t st);

// caller vector (reference layout, HR; host or device memory) -> device SoA vector r of v
template <typename HR, typename VR> inline void upload_vec(const void *h, Vec<VR> v, int r, DevBuf &stage, hipStream_t st)
{
    const size_t n = (size_t)6 * v.Vh;
    if (stage.bytes() < sizeof(HR) * n)
        stage.alloc(sizeof(HR) * n);
    copy_from_host(stage.as<void>(), h, sizeof(HR) * n, st);
    vec_host_to_soa<HR, VR>(stage.as<HR>(), v, r, st);
    KQ_HIP_CHECK(hipStreamSynchronize(st));
}

template <typename VR, typename HR> inline void download_vec(Vec<VR> v, int r, void *h, DevBuf &stage, hipStream_t st)
{
    const size_t n = (size_t)6 * v.Vh;
    if (stage.bytes() < sizeof(HR) * n)
        stage.alloc(sizeof(HR) * n);
    vec_soa_to_host<VR, HR>(v, r, stage.as<HR>(), st);
    copy_to_host(h, stage.as<void>(), sizeof(HR) * n, st);
}

// autotuned (nr, nt) for the hopping term
std::pair<int, int> tuned_launch(int slot, int mode, int nrhs);
// occupancy cap (waves per SIMD, 0 = none) measured best for (nr, nt) by tuned_launch
int tuned_cap(int slot, int mode, int nr, int nt);

}  // namespace kq
