/*
 * Oracle for the KuQCD HISQ even-odd multishift CG path: a plain-C CPU
 * restatement of the reference algorithm.
 *
 * TEST INFRASTRUCTURE ONLY. Only the test suite, the smoke check and the
 * benchmark's CPU-baseline leg load this library, and only as the checker or
 * the timed CPU baseline; the product library never links or calls it.
 *
 * Built with -ffp-contract=off so that every a*b+c is two IEEE roundings, as
 * in the reference's x86 build (no -march, no -ffast-math).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define KQO_BLK 4096

/* Checkerboard coordinates of site s of a parity: s = xcb + (Lx/2)*(y + Ly*(z + Lz*t))
 * (index_utils.cpp:17-33 applied to the checkerboard lattice of latt_base.cpp:119-123). */
static inline void kqo_cb_coords(const int L[4], long s, int c[4])
{
    const long lxh = L[0] / 2;
    long r = s / lxh;
    c[0] = (int)(s - r * lxh);
    c[1] = (int)(r % L[1]);
    r /= L[1];
    c[2] = (int)(r % L[2]);
    c[3] = (int)(r / L[2]);
}

/* Checkerboard index of the neighbour of site s (parity q) at hop h along mu.
 * Follows DiracHISQ::InitNbIdxParity (dirac_hisq.cpp:139-231):
 *   full x = 2*xcb + ((y+z+t+q)&1)              (index_utils.cpp:91-100)
 *   coordinate mu += h, wrapped into [0, L_mu)  (space_time.cpp:187-197)
 *   x -> x>>1, index x + Lxh*(y + Ly*(z + Lz*t)) (index_utils.cpp:102-111, space_time.cpp:204-211)
 * The same formula defines the operator for odd or unit extents. */
long kqo_nb(const int L[4], int q, long s, int mu, int h)
{
    int c[4];
    kqo_cb_coords(L, s, c);
    c[0] = 2 * c[0] + ((c[1] + c[2] + c[3] + q) & 1);
    int v = (c[mu] + h) % L[mu];
    while (v < 0)
        v += L[mu];
    c[mu] = v;
    const long lxh = L[0] / 2;
    return (long)(c[0] >> 1) + lxh * (c[1] + (long)L[1] * (c[2] + (long)L[2] * c[3]));
}

/* Full-lattice coordinates of checkerboard site s of parity q. */
void kqo_full_coords(const int L[4], int q, long s, int out[4])
{
    kqo_cb_coords(L, s, out);
    out[0] = 2 * out[0] + ((out[1] + out[2] + out[3] + q) & 1);
}

int kqo_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void kqo_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0)
        omp_set_num_threads(n);
#else
    (void)n;
#endif
}

#define REAL double
#define SFX _d
#include "kq_oracle_t.h"
#undef REAL
#undef SFX

#define REAL float
#define SFX _f
#include "kq_oracle_t.h"
#undef REAL
#undef SFX
