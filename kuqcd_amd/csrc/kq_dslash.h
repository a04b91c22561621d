// Host-side interface of the HISQ even-odd hopping-term kernels.
#pragma once

#include "kq_fields.h"

namespace kq {

enum DslashEpilogue {
    EPI_STORE = 0,     // dst = D src
    EPI_AXMY_DOT = 1,  // dst = sigma[r] * aux - D src ; partial[r][blk] = Re<aux, dst>
};

template <class LF, class LL, typename SR> struct DslashArgs {
    using R = typename LF::R;
    using V2 = c2_t<SR>;
    Geom g;
    int dst_parity;
    LF fat;
    LL lng;
    Vec<SR> src, dst, aux;
    const double *sigma;  // per right-hand side (EPI_AXMY_DOT)
    double *partial;      // [nrhs_total][nblk] (EPI_AXMY_DOT)
    const int *run;       // device flag; kernel does nothing when *run == 0 (may be null)
    const int *rlist;     // compact list of active right-hand sides (may be null = identity)
    const int *nact;      // number of entries of rlist in use (may be null = nrhs)
    int nrhs;             // right-hand sides addressed by this launch
    uint32_t nblk;        // grid.x, the partial-sum stride
};

// block sizes launch_dslash accepts: 64, 128 and 256 threads
constexpr int kMinDslashNT = 64;
// Block size of the hopping terms whose partial sums feed the reference method, the plain
// FP64 solve and the FP64 verification. One partial per workgroup, added in a fixed order,
// makes a reduction deterministic for a given block size; fixing the block size (instead of
// taking the autotuner's timing-dependent pick) makes these results bitwise reproducible from
// one process to the next.
constexpr int kReproNT = 256;

// Launch D over all sites of parity dst_parity for every right-hand side.
// nr = right-hand sides handled per thread (1, 2, 3, 4 or 8), nt = threads per block,
// cap = at most this many wavefronts per SIMD (0: as many as the registers allow). The
// cap is imposed with unused dynamic LDS; fewer resident waves keep the gathered
// neighbour vectors of the batched terms in L2 (measured up to 19 % faster at 32^4).
template <class LF, class LL, typename SR>
void launch_dslash(DslashArgs<LF, LL, SR> a, int epi, int nr, int nt, hipStream_t st, int cap = 0);

inline uint32_t dslash_nblk(uint32_t Vh, int nt)
{
    uint32_t nb = (Vh + nt - 1) / nt;
    return nb;
}

// Row length Lx/2 of the lattices whose batched hopping terms stage rows in LDS (Lx = 32, as at
// 32^4): a compile-time row length gives the staged kernel a static LDS size
constexpr int kStageLxh = 16;

// KQ_STAGE=0 turns the LDS row staging of the batched hopping terms off (a tuning switch;
// results are bitwise the same either way)
inline bool stage_enabled()
{
    static const bool on = [] {
        const char *e = getenv("KQ_STAGE");
        return !(e && e[0] == '0');
    }();
    return on;
}

// Whether launch_dslash runs the LDS-staged kernel for nr right-hand sides per thread and block
// size nt on this lattice (its NR * 3 * 352 staged colour entries, at most 67.6 KB for eight
// FP32 or 33.8 KB for two FP64 right-hand sides, fit a workgroup's LDS on gfx950)
inline bool dslash_staged(const Geom &g, int nr, int nt)
{
    return nr >= 2 && nt == 256 && g.lxh == kStageLxh && stage_enabled() && stage_rows(g, nt) > 0;
}

}  // namespace kq
