"""Oracle for the HISQ even-odd multishift CG path (TEST INFRASTRUCTURE ONLY).

A plain-C restatement of the reference algorithm (kq_oracle.c, built by
oracle/Makefile into liboracle) with numpy helpers. Only the test suite, the
smoke check and the benchmark's CPU-baseline leg may use it, and only as the
checker or as the timed CPU baseline; the product (kuqcd_amd) never imports it.

Pinning (tests/test_oracle.py): the reference test's known answer
(test/test_solver.cpp:74-123), closed forms for constant and free fields,
anti-Hermiticity, gauge covariance, an independent dense-matrix construction
on small lattices, and dense linear solves of the shifted systems.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libkq_oracle.so"
_lib = None


def build() -> None:
    r = subprocess.run(["make", "-s", "-C", str(HERE)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        l = C.CDLL(str(LIB))
        I4 = C.POINTER(C.c_int)
        VP = C.c_void_p
        l.kqo_nb.restype = C.c_long
        l.kqo_nb.argtypes = [I4, C.c_int, C.c_long, C.c_int, C.c_int]
        l.kqo_full_coords.restype = None
        l.kqo_full_coords.argtypes = [I4, C.c_int, C.c_long, I4]
        for sfx, R in (("_d", C.c_double), ("_f", C.c_float)):
            getattr(l, "kqo_dslash" + sfx).argtypes = [VP, VP, VP, VP, I4, C.c_int]
            getattr(l, "kqo_dslash" + sfx).restype = None
            getattr(l, "kqo_mdaggm" + sfx).argtypes = [VP, VP, VP, VP, VP, I4, C.c_int]
            getattr(l, "kqo_mdaggm" + sfx).restype = None
            getattr(l, "kqo_cgm" + sfx).argtypes = [VP, VP, I4, C.c_int, C.c_int, VP, VP, C.POINTER(VP), C.c_double,
                                                   C.c_int]
            getattr(l, "kqo_cgm" + sfx).restype = C.c_int
            getattr(l, "kqo_realdot" + sfx).argtypes = [VP, VP, C.c_long]
            getattr(l, "kqo_realdot" + sfx).restype = R
        l.kqo_num_threads.restype = C.c_int
        l.kqo_set_threads.argtypes = [C.c_int]
        l.kqo_set_threads.restype = None
        _lib = l
    return _lib


def _i4(lat):
    return (C.c_int * 4)(*[int(v) for v in lat])


def _p(a):
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


def _sfx(dtype):
    return "_d" if np.dtype(dtype) == np.float64 else "_f"


def nb(lat, q, s, mu, h) -> int:
    return lib().kqo_nb(_i4(lat), q, s, mu, h)


def full_coords(lat, q, s):
    out = (C.c_int * 4)()
    lib().kqo_full_coords(_i4(lat), q, s, out)
    return tuple(out)


def dslash(fat, lng, src, lat, dst_parity):
    """dst = D src (dirac_hisq.cpp:409-502) in the dtype of the inputs."""
    dst = np.empty_like(src)
    getattr(lib(), "kqo_dslash" + _sfx(src.dtype))(_p(fat), _p(lng), _p(src), _p(dst), _i4(lat), dst_parity)
    return dst


def mdaggm(fat, lng, src, lat, parity):
    """D D src (dirac_hisq.cpp:457-462)."""
    dst = np.empty_like(src)
    tmp = np.empty_like(src)
    getattr(lib(), "kqo_mdaggm" + _sfx(src.dtype))(_p(fat), _p(lng), _p(src), _p(dst), _p(tmp), _i4(lat), parity)
    return dst


def cgm(fat, lng, lat, parity, shifts, b, tol, max_iter):
    """CGM::Invert (inverter_cgm.cpp:17-87) in the dtype of b; returns (x list, iterations).
    parity 0 or 1: the even-odd operator on fields of that parity; parity -1: the full operator
    (DiracHISQ with KUQCD_HISQ_DIRAC) on fields of both parities, even block first."""
    dt = b.dtype
    sh = np.ascontiguousarray(shifts, dtype=dt)
    xs = [np.empty_like(b) for _ in range(len(sh))]
    xp = (C.c_void_p * len(xs))(*[x.ctypes.data for x in xs])
    it = getattr(lib(), "kqo_cgm" + _sfx(dt))(_p(fat), _p(lng), _i4(lat), parity, len(sh), _p(sh), _p(b), xp,
                                              float(tol), int(max_iter))
    return xs, it


def set_threads(n: int) -> None:
    lib().kqo_set_threads(n)


def num_threads() -> int:
    return lib().kqo_num_threads()
