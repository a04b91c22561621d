"""ctypes binding of libkuqcd.so (the C ABI declared in include/kuqcd_hip.h).

The library is required: importing the bindings raises if the in-tree
libkuqcd.so has not been built. There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libkuqcd.so"

# ---------------------------------------------------------------- enums (kuqcd_enum.h)
KUQCD_SUCCESS = 0
KUQCD_ERR_INVALID_PARAM = -1
KUQCD_ERR_UNSUPPORTED = -2
KUQCD_ERR_NULL_POINTER = -3
KUQCD_ERR_NO_MEMORY = -4
KUQCD_ERR_UNKNOWN_TYPE = -5
KUQCD_ERR_UNEXPECTED_CALL_ORDER = -6
KUQCD_ERR_NO_RESOURCE = -7
KUQCD_ERR_UNKNOWN_ERROR = -8

KUQCD_EVEN, KUQCD_ODD, KUQCD_ALL, KUQCD_INVALID_PARITY = 0, 1, 2, 3
KUQCD_HALF, KUQCD_SINGLE, KUQCD_DOUBLE, KUQCD_INVALID_PRECISION = 2, 4, 8, 9
KUQCD_XYZT_ORDER, KUQCD_TZYX_ORDER = 0, 1
KUQCD_NORMAL_ORDER, KUQCD_BATCH_ORDER, KUQCD_INVALID_ORDER = 0, 1, 2
KUQCD_EVEN_EVEN, KUQCD_ODD_ODD = 0, 1
KUQCD_WILSON_LINK, KUQCD_FAT_LINK, KUQCD_LONG_LINK = 0, 1, 2
KUQCD_GAUGE_RECON_NO, KUQCD_GAUGE_RECON_U14, KUQCD_GAUGE_RECON_12, KUQCD_GAUGE_RECON_8 = 18, 14, 12, 8
KUQCD_INVALID_RECON = 0
KUQCD_WILSON_DIRAC, KUQCD_WILSONPC_DIRAC, KUQCD_HISQ_DIRAC, KUQCD_HISQPC_DIRAC, KUQCD_INVALID_DIRAC = range(5)
KUQCD_CG, KUQCD_CGM, KUQCD_BICGSTAB, KUQCD_BICGSTABM, KUQCD_INVALID_INVERTER = range(5)

KQ_MODE_REFERENCE, KQ_MODE_FP64, KQ_MODE_MIXED_FP32, KQ_MODE_MIXED_FP16, KQ_MODE_MIXED_FP32_L16 = 0, 1, 2, 3, 4
KQ_MAX_SHIFT = 32
KQ_GEN_HAAR, KQ_GEN_WEAK = 0, 1


class KuQCDGaugeParam(C.Structure):
    _fields_ = [
        ("lat", C.c_uint * 4),
        ("halo_depth", C.c_int),
        ("od", C.c_int),
        ("site_subset", C.c_int),
        ("gauge_order", C.c_int),
        ("prec", C.c_int),
        ("link_type", C.c_int),
        ("recon", C.c_int),
    ]


class KuQCDInvertParam(C.Structure):
    _fields_ = [
        ("dslash_type", C.c_int),
        ("inverter_type", C.c_int),
        ("matpc_type", C.c_int),
        ("prec", C.c_int),
        ("num_shift", C.c_int),
        ("max_iter", C.c_int),
        ("iter", C.c_int),
        ("tol", C.c_double),
        ("shifts", C.POINTER(C.c_double)),
    ]


class KqSolveOptions(C.Structure):
    _fields_ = [
        ("mode", C.c_int),
        ("reliable_delta", C.c_double),
        ("refine", C.c_int),
        ("rel_tol", C.c_double),
        ("learned", C.c_int),
        ("chunk", C.c_int),
        ("nr", C.c_int),
        ("nt", C.c_int),
        ("verbose", C.c_int),
        ("fork_tol", C.c_double),
        ("learned_tol", C.c_double),
    ]


class KqSolveReport(C.Structure):
    _fields_ = [
        ("iter", C.c_int),
        ("reliable_updates", C.c_int),
        ("refine_rounds", C.c_int),
        ("refine_iter", C.c_int * KQ_MAX_SHIFT),
        ("true_rel_res", C.c_double * KQ_MAX_SHIFT),
        ("seconds", C.c_double),
        ("seconds_multishift", C.c_double),
        ("seconds_refine", C.c_double),
        ("learned_used", C.c_int),
        ("learned_fallback", C.c_int),
        ("initial_rel_res", C.c_double * KQ_MAX_SHIFT),
        ("seconds_verify", C.c_double),
    ]


class KqBenchResult(C.Structure):
    _fields_ = [
        ("ms", C.c_double),
        ("bytes", C.c_double),
        ("gbs", C.c_double),
        ("gflops", C.c_double),
        ("hbm_frac", C.c_double),
        ("nr", C.c_int),
        ("nt", C.c_int),
    ]


class KqPtWeights(C.Structure):
    _fields_ = [
        ("w1", ((C.c_float * 2) * 16) * 4),
        ("w2", ((C.c_float * 2) * 16) * 4),
        ("w0", C.c_float * 2),
    ]


KQ_TRAIN_HIST = 64
KQ_PT_PARAMS = 2 * 4 * 16 + 1  # complex weights


class KqTrainOptions(C.Structure):
    _fields_ = [
        ("steps", C.c_int),
        ("batch", C.c_int),
        ("lr", C.c_double),
        ("lr_final", C.c_double),
        ("seed", C.c_uint64),
        ("init", C.c_int),
        ("init_scale", C.c_double),
        ("verbose", C.c_int),
    ]


class KqTrainReport(C.Structure):
    _fields_ = [
        ("steps", C.c_int),
        ("nsamples", C.c_int),
        ("ngauges", C.c_int),
        ("loss_initial", C.c_double),
        ("loss_final", C.c_double),
        ("loss_hist", C.c_double * KQ_TRAIN_HIST),
        ("seconds", C.c_double),
    ]


class KqRefineOptions(C.Structure):
    _fields_ = [
        ("rel_tol", C.c_double),
        ("learned", C.c_int),
        ("mode", C.c_int),
        ("net_format", C.c_int),
        ("max_iter", C.c_int),
        ("delta", C.c_double),
        ("chunk", C.c_int),
        ("stall_iter", C.c_int),
        ("verbose", C.c_int),
    ]


class KqRefineReport(C.Structure):
    _fields_ = [
        ("iter", C.c_int),
        ("iter_learned", C.c_int),
        ("reliable_updates", C.c_int),
        ("fallback", C.c_int),
        ("converged", C.c_int),
        ("rel_res_initial", C.c_double),
        ("true_rel_res", C.c_double),
        ("seconds", C.c_double),
    ]


_VP = C.c_void_p
_PROTOS = {
    "kq_init": (C.c_int, []),
    "kq_init_grid": (C.c_int, [C.POINTER(C.c_int)]),
    "kq_load_gaugefield": (C.c_int, [_VP, C.POINTER(KuQCDGaugeParam)]),
    "kq_solve_multishift": (C.c_int, [C.POINTER(_VP), _VP, C.POINTER(KuQCDInvertParam)]),
    "kq_gauge_free": (None, []),
    "kq_end": (None, []),
    "kq_default_options": (None, [C.POINTER(KqSolveOptions)]),
    "kq_solve_multishift_ex": (C.c_int, [C.POINTER(_VP), _VP, C.POINTER(KuQCDInvertParam),
                                         C.POINTER(KqSolveOptions), C.POINTER(KqSolveReport)]),
    "kq_solve_multishift_batched": (C.c_int, [C.c_int, C.POINTER(_VP), C.POINTER(_VP), C.POINTER(KuQCDInvertParam),
                                              C.POINTER(KqSolveOptions), C.POINTER(KqSolveReport)]),
    "kq_dslash": (C.c_int, [_VP, _VP, C.c_int, C.c_int]),
    "kq_dslash_batched": (C.c_int, [C.c_int, _VP, _VP, C.c_int, C.c_int, C.c_int, C.c_int]),
    "kq_mdaggm": (C.c_int, [_VP, _VP, C.c_int, C.c_double, C.c_int]),
    "kq_mdaggm_noalloc": (C.c_int, [_VP, _VP, C.c_int, C.c_double]),
    "kq_bench_dslash": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(KqBenchResult)]),
    "kq_bench_copy": (C.c_int, [C.c_size_t, C.c_int, C.POINTER(KqBenchResult)]),
    "kq_autotune_dslash": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "kq_gen_gauge": (C.c_int, [_VP, _VP, C.POINTER(C.c_uint), C.c_int, C.c_double, C.c_double, C.c_double,
                               C.c_uint64, C.c_uint64, C.c_int]),
    "kq_gen_source": (C.c_int, [_VP, C.POINTER(C.c_uint), C.c_int, C.c_uint64]),
    "kq_device_alloc": (_VP, [C.c_size_t]),
    "kq_device_free": (None, [_VP]),
    "kq_memcpy": (C.c_int, [_VP, _VP, C.c_size_t]),
    "kq_synchronize": (C.c_int, []),
    "kq_ptnet_init_weights": (None, [C.POINTER(KqPtWeights), C.c_uint64, C.c_double]),
    "kq_ptnet_set_weights": (C.c_int, [C.POINTER(KqPtWeights)]),
    "kq_ptnet_get_weights": (C.c_int, [C.POINTER(KqPtWeights)]),
    "kq_ptnet_norms": (C.c_int, [C.POINTER(C.c_double)]),
    "kq_ptnet_forward": (C.c_int, [C.POINTER(KqPtWeights), C.c_int, C.c_int, _VP, _VP, _VP]),
    "kq_ptnet_forward_h16": (C.c_int, [C.POINTER(KqPtWeights), C.c_int, C.c_int, _VP, _VP, _VP]),
    "kq_ptnet_loss_grad": (C.c_int, [C.POINTER(KqPtWeights), C.c_int, C.c_double, _VP, _VP,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "kq_learned_collect": (C.c_int, [C.c_int, C.c_double, C.c_int, C.c_uint64, C.c_double, C.c_double,
                                     C.POINTER(C.c_int)]),
    "kq_learned_clear": (None, []),
    "kq_learned_train": (C.c_int, [C.POINTER(KqTrainOptions), C.POINTER(KqTrainReport)]),
    "kq_learned_eval": (C.c_int, [C.POINTER(KqPtWeights), C.POINTER(C.c_double)]),
    "kq_default_refine_options": (None, [C.POINTER(KqRefineOptions)]),
    "kq_fcg_refine": (C.c_int, [C.c_int, C.c_double, _VP, _VP, C.POINTER(KqRefineOptions),
                                C.POINTER(KqRefineReport)]),
    "kq_last_error": (C.c_char_p, []),
    "kq_set_verbosity": (None, [C.c_int]),
    "kq_device_count": (C.c_int, []),
    "kq_selftest_fastdiv": (C.c_long, [C.c_uint32, C.c_uint32]),
    "kq_selftest_fold_keeps": (C.c_int, [C.c_int, C.c_int, C.c_int]),
    "kq_selftest_seg_coef": (None, [_VP, C.c_int, C.c_int, C.c_int, _VP]),
    "kq_neighbour_index": (C.c_long, [C.POINTER(C.c_int), C.c_int, C.c_long, C.c_int, C.c_int]),
}

_lib = None


def load(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load libkuqcd.so (built in-tree by kuqcd_amd.build). Raises if absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build it with `python -m kuqcd_amd.build` "
            "(the HIP library is required; there is no CPU fallback)")
    lib = C.CDLL(str(p))
    for name, (res, args) in _PROTOS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib


def lib() -> C.CDLL:
    return load()


def last_error() -> str:
    s = lib().kq_last_error()
    return s.decode() if s else ""
