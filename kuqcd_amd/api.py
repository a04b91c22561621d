"""Python mirror of the KuQCD API (kuqcd.h:29-34) over the C ABI of libkuqcd.so.

Same function names, argument meaning and status codes as the reference:
each call returns a KuQCDStatus int, arrays are numpy arrays in the
reference's host layouts (kuqcd.cpp:103-169):

  gauge  : shape (V, 4, 9, 2), i = parity*Vh + checkerboard index, l = 3*row + col
  vector : shape (Vh, 3, 2) on the sites of one parity
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from ._lib import *  # noqa: F401,F403  (enums and structs)


def _ptr(a: np.ndarray | None) -> C.c_void_p:
    if a is None:
        return C.c_void_p(None)
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return C.c_void_p(a.ctypes.data)


def volume(lat) -> int:
    return int(lat[0]) * int(lat[1]) * int(lat[2]) * int(lat[3])


# ---------------------------------------------------------------- reference API
def kuqcd_init(grid=(1, 1, 1, 1)) -> int:
    """kuqcd_init (kuqcd.cpp:288-301) with the grid given directly.
    The stdin form is kuqcd_init_stdin()."""
    g = (C.c_int * 4)(*grid)
    return L.lib().kq_init_grid(g)


def kuqcd_init_stdin() -> int:
    """kuqcd_init reading gridx=.. lines from the process's stdin."""
    return L.lib().kq_init()


def gauge_param(lat, prec=L.KUQCD_DOUBLE, link_type=L.KUQCD_FAT_LINK, order=L.KUQCD_NORMAL_ORDER,
                halo_depth=3, site_subset=L.KUQCD_ALL, od=L.KUQCD_TZYX_ORDER,
                recon=L.KUQCD_GAUGE_RECON_NO) -> L.KuQCDGaugeParam:
    p = L.KuQCDGaugeParam()
    for i in range(4):
        p.lat[i] = int(lat[i])
    p.halo_depth = halo_depth
    p.od = od
    p.site_subset = site_subset
    p.gauge_order = order
    p.prec = prec
    p.link_type = link_type
    p.recon = recon
    return p


def kuqcd_load_gaugefield(h_gauge: np.ndarray | None, param: L.KuQCDGaugeParam | None) -> int:
    return L.lib().kq_load_gaugefield(_ptr(h_gauge), C.byref(param) if param is not None else None)


@dataclass
class InvertParam:
    """Owner of a KuQCDInvertParam and the shift array it points to."""
    shifts: np.ndarray
    tol: float
    max_iter: int = 10000
    prec: int = L.KUQCD_DOUBLE
    matpc_type: int = L.KUQCD_EVEN_EVEN
    dslash_type: int = L.KUQCD_HISQPC_DIRAC
    inverter_type: int = L.KUQCD_CGM
    num_shift: int | None = None
    c: L.KuQCDInvertParam = field(default_factory=L.KuQCDInvertParam)

    def struct(self) -> L.KuQCDInvertParam:
        self.shifts = np.ascontiguousarray(self.shifts, dtype=np.float64)
        c = self.c
        c.dslash_type = self.dslash_type
        c.inverter_type = self.inverter_type
        c.matpc_type = self.matpc_type
        c.prec = self.prec
        c.num_shift = self.num_shift if self.num_shift is not None else int(self.shifts.shape[-1])
        c.max_iter = self.max_iter
        c.tol = self.tol
        c.shifts = self.shifts.ctypes.data_as(C.POINTER(C.c_double))
        return c

    @property
    def iter(self) -> int:
        return self.c.iter


def kuqcd_solve_multishift(x: list, b: np.ndarray | None, param: InvertParam | None) -> int:
    """kuqcd_solve_multishift (kuqcd.cpp:404-445): x[k] <- solution of shift k."""
    if x is None:
        xp = None
    else:
        xp = (C.c_void_p * max(1, len(x)))(*[_ptr(a).value for a in x])
    return L.lib().kq_solve_multishift(xp, _ptr(b), C.byref(param.struct()) if param is not None else None)


def kuqcd_gauge_free() -> None:
    L.lib().kq_gauge_free()


def kuqcd_end() -> None:
    L.lib().kq_end()


# ---------------------------------------------------------------- extensions
def default_options(**kw) -> L.KqSolveOptions:
    o = L.KqSolveOptions()
    L.lib().kq_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def solve_ex(x: list, b: np.ndarray, param: InvertParam, opt: L.KqSolveOptions | None = None):
    rep = L.KqSolveReport()
    xp = (C.c_void_p * len(x))(*[_ptr(a).value for a in x])
    st = L.lib().kq_solve_multishift_ex(xp, _ptr(b), C.byref(param.struct()),
                                        C.byref(opt) if opt is not None else None, C.byref(rep))
    return st, rep


def solve_batched(x: list, b: list, param: InvertParam, opt: L.KqSolveOptions | None = None):
    """x: nrhs*num_shift arrays (rhs-major), b: nrhs arrays; param.shifts shape (nrhs, num_shift)."""
    nrhs = len(b)
    reps = (L.KqSolveReport * nrhs)()
    xp = (C.c_void_p * len(x))(*[_ptr(a).value for a in x])
    bp = (C.c_void_p * nrhs)(*[_ptr(a).value for a in b])
    param.num_shift = int(np.asarray(param.shifts).shape[-1])
    st = L.lib().kq_solve_multishift_batched(nrhs, xp, bp, C.byref(param.struct()),
                                             C.byref(opt) if opt is not None else None, reps)
    return st, list(reps)


def dslash(src: np.ndarray, dst: np.ndarray, dst_parity: int, mode: int = L.KQ_MODE_REFERENCE) -> int:
    return L.lib().kq_dslash(_ptr(src), _ptr(dst), dst_parity, mode)


def dslash_batched(src: np.ndarray, dst: np.ndarray, dst_parity: int, mode: int, nr: int, nt: int) -> int:
    return L.lib().kq_dslash_batched(int(src.shape[0]), _ptr(src), _ptr(dst), dst_parity, mode, nr, nt)


def mdaggm(src: np.ndarray, dst: np.ndarray, parity: int, sigma: float, mode: int = L.KQ_MODE_REFERENCE) -> int:
    return L.lib().kq_mdaggm(_ptr(src), _ptr(dst), parity, sigma, mode)


def bench_dslash(mode: int, nrhs: int = 1, nr: int = 0, nt: int = 0, iters: int = 20):
    r = L.KqBenchResult()
    st = L.lib().kq_bench_dslash(mode, nrhs, nr, nt, iters, C.byref(r))
    return st, r


def bench_copy(nbytes: int, iters: int = 20):
    r = L.KqBenchResult()
    st = L.lib().kq_bench_copy(nbytes, iters, C.byref(r))
    return st, r


def gen_gauge(lat, kind=L.KQ_GEN_HAAR, eps=0.3, fat_scale=0.9, c_naik=-1.0 / 24.0, seed_fat=1, seed_long=2,
              phases=1, fat=True, long=True):
    """Synthetic fat/long links in the reference host layout, shape (V, 4, 9, 2) float64."""
    V = volume(lat)
    f = np.empty((V, 4, 9, 2), dtype=np.float64) if fat else None
    g = np.empty((V, 4, 9, 2), dtype=np.float64) if long else None
    la = (C.c_uint * 4)(*[int(v) for v in lat])
    rc = L.lib().kq_gen_gauge(_ptr(f), _ptr(g), la, kind, eps, fat_scale, c_naik, seed_fat, seed_long, phases)
    if rc != 0:
        raise ValueError("kq_gen_gauge failed (invalid lattice)")
    return f, g


def gen_source(lat, parity=0, seed=3) -> np.ndarray:
    Vh = volume(lat) // 2
    b = np.empty((Vh, 3, 2), dtype=np.float64)
    la = (C.c_uint * 4)(*[int(v) for v in lat])
    if L.lib().kq_gen_source(_ptr(b), la, parity, seed) != 0:
        raise ValueError("kq_gen_source failed")
    return b


def neighbour_index(lat, q: int, s: int, mu: int, h: int) -> int:
    la = (C.c_int * 4)(*[int(v) for v in lat])
    return L.lib().kq_neighbour_index(la, q, s, mu, h)


def load_gauge_pair(fat: np.ndarray, lng: np.ndarray, lat, prec=L.KUQCD_DOUBLE, order=L.KUQCD_NORMAL_ORDER) -> None:
    """Load fat then long links (the reference call order); raises on error."""
    dt = np.float64 if prec == L.KUQCD_DOUBLE else np.float32
    for arr, lt in ((fat, L.KUQCD_FAT_LINK), (lng, L.KUQCD_LONG_LINK)):
        a = np.ascontiguousarray(arr, dtype=dt)
        st = kuqcd_load_gaugefield(a, gauge_param(lat, prec=prec, link_type=lt, order=order))
        if st != L.KUQCD_SUCCESS:
            raise RuntimeError(f"kuqcd_load_gaugefield failed: {st} {L.last_error()}")


def set_verbosity(v: int) -> None:
    L.lib().kq_set_verbosity(v)


def device_count() -> int:
    return L.lib().kq_device_count()
