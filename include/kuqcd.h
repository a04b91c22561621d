/*
 * KuQCD public API, MI355X implementation (drop-in for the reference
 * include/kuqcd.h:26-34): the same five C++ functions in namespace kuqcd,
 * exported from libkuqcd.so (SONAME libkuqcd.so.1).
 *
 *   init -> load FAT -> load LONG -> solve* -> gauge_free -> end
 *
 * Like the reference header (kuqcd.h:18-22) it also declares the inner classes
 * (Lattice, LattBase, SpinorField, GaugeField, DiracHISQ, CGM and the BLAS-1
 * functions: kuqcd_classes.h), here over fields in GPU memory.
 *
 * The reference header also pulls in <mpi.h> and <omp.h>; this one includes
 * <mpi.h> only when the caller's toolchain has it, since the library runs one
 * process per GPU and does not call MPI itself.
 */
#ifndef KUQCD_H
#define KUQCD_H

#include <cstdio>
#include "kuqcd_enum.h"
#include "kuqcd_classes.h"

#if defined(__has_include)
#if __has_include(<mpi.h>)
#include <mpi.h>
#endif
#endif

#define kuqcd_export __attribute__((visibility("default")))

namespace kuqcd {

/* kuqcd.h:29 -- reads gridx=..gridt= lines from stdin (command_line.cpp:31-84);
 * the product of the grid must be 1 (one GPU per process). */
kuqcd_export KuQCDStatus kuqcd_init();
/* kuqcd.h:30 -- copies the host links (reference host layout) to the GPU */
kuqcd_export KuQCDStatus kuqcd_load_gaugefield(void *h_gauge, KuQCDGaugeParam *param);
/* kuqcd.h:31 -- even-odd HISQ multishift CG (CGM) on the GPU */
kuqcd_export KuQCDStatus kuqcd_solve_multishift(void **_hp_x, void *_hp_b, KuQCDInvertParam *param);
/* kuqcd.h:33 */
kuqcd_export void kuqcd_gauge_free();
/* kuqcd.h:34 */
kuqcd_export void kuqcd_end();

}  // namespace kuqcd

#endif
