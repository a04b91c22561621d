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


def kq_synchronize() -> int:
    """Wait for all work the library queued on its stream (0 on success)."""
    return L.lib().kq_synchronize()


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


# ---------------------------------------------------------------- learned preconditioner
def weights_to_array(w: L.KqPtWeights) -> tuple[np.ndarray, np.ndarray, complex]:
    """(W1 [4,16] complex, W2 [4,16] complex, W0) of a KqPtWeights struct."""
    a = np.frombuffer(bytes(w), dtype=np.float32).astype(np.float64)
    c = a[0::2] + 1j * a[1::2]
    return c[:64].reshape(4, 16), c[64:128].reshape(4, 16), complex(c[128])


def weights_from_array(w1, w2, w0) -> L.KqPtWeights:
    c = np.concatenate([np.asarray(w1, np.complex128).ravel(), np.asarray(w2, np.complex128).ravel(),
                        np.asarray([w0], np.complex128)])
    f = np.empty(2 * c.size, dtype=np.float32)
    f[0::2], f[1::2] = c.real, c.imag
    w = L.KqPtWeights()
    C.memmove(C.addressof(w), f.ctypes.data, C.sizeof(w))
    return w


def ptnet_init_weights(seed: int = 1, scale: float = 0.05) -> L.KqPtWeights:
    w = L.KqPtWeights()
    L.lib().kq_ptnet_init_weights(C.byref(w), seed, scale)
    return w


def ptnet_set_weights(w: L.KqPtWeights) -> int:
    return L.lib().kq_ptnet_set_weights(C.byref(w))


def ptnet_get_weights() -> L.KqPtWeights:
    w = L.KqPtWeights()
    st = L.lib().kq_ptnet_get_weights(C.byref(w))
    if st != L.KUQCD_SUCCESS:
        raise RuntimeError(f"kq_ptnet_get_weights: {st} {L.last_error()}")
    return w


def ptnet_norms() -> tuple[float, float]:
    nu = (C.c_double * 2)()
    st = L.lib().kq_ptnet_norms(nu)
    if st != L.KUQCD_SUCCESS:
        raise RuntimeError(f"kq_ptnet_norms: {st} {L.last_error()}")
    return nu[0], nu[1]


def ptnet_forward(w: L.KqPtWeights, r: np.ndarray, parity: int = 0, link_format: int = L.KQ_MODE_MIXED_FP32,
                  hidden: bool = False, hidden16: bool = False):
    """z = M r (and the hidden channels, shape (4, Vh, 3, 2), if hidden). hidden16: the channels pass
    between the layers as bfloat16, as in the flexible CG (kq_ptnet_forward_h16)."""
    z = np.empty_like(r)
    h = np.empty((4,) + r.shape, dtype=np.float64) if hidden else None
    fn = L.lib().kq_ptnet_forward_h16 if hidden16 else L.lib().kq_ptnet_forward
    st = fn(C.byref(w), parity, link_format, _ptr(r), _ptr(z), _ptr(h))
    if st != L.KUQCD_SUCCESS:
        raise RuntimeError(f"kq_ptnet_forward: {st} {L.last_error()}")
    return (z, h) if hidden else z


def ptnet_loss_grad(w: L.KqPtWeights, r: np.ndarray, e: np.ndarray, sigma: float, parity: int = 0):
    """(loss, dL/dconj(W) as complex [129]: W1 (64), W2 (64), W0)."""
    loss = C.c_double()
    g = np.empty(2 * L.KQ_PT_PARAMS, dtype=np.float64)
    st = L.lib().kq_ptnet_loss_grad(C.byref(w), parity, sigma, _ptr(r), _ptr(e), C.byref(loss),
                                    g.ctypes.data_as(C.POINTER(C.c_double)))
    if st != L.KUQCD_SUCCESS:
        raise RuntimeError(f"kq_ptnet_loss_grad: {st} {L.last_error()}")
    return loss.value, g[0::2] + 1j * g[1::2]


def learned_collect(parity: int, sigma: float, nsrc: int = 1, seed: int = 1, rel_tol: float = 1e-11,
                    min_level: float = 1e-8) -> int:
    n = C.c_int()
    st = L.lib().kq_learned_collect(parity, sigma, nsrc, seed, rel_tol, min_level, C.byref(n))
    if st != L.KUQCD_SUCCESS:
        raise RuntimeError(f"kq_learned_collect: {st} {L.last_error()}")
    return n.value


def learned_clear() -> None:
    L.lib().kq_learned_clear()


def learned_train(steps=200, batch=8, lr=3e-4, lr_final=3e-5, seed=1, init=2, init_scale=0.01,
                  verbose=0) -> L.KqTrainReport:
    o = L.KqTrainOptions(steps=steps, batch=batch, lr=lr, lr_final=lr_final, seed=seed, init=init,
                         init_scale=init_scale, verbose=verbose)
    rep = L.KqTrainReport()
    st = L.lib().kq_learned_train(C.byref(o), C.byref(rep))
    if st != L.KUQCD_SUCCESS:
        raise RuntimeError(f"kq_learned_train: {st} {L.last_error()}")
    return rep


def learned_eval(w: L.KqPtWeights) -> float:
    m = C.c_double()
    st = L.lib().kq_learned_eval(C.byref(w), C.byref(m))
    if st != L.KUQCD_SUCCESS:
        raise RuntimeError(f"kq_learned_eval: {st} {L.last_error()}")
    return m.value


def refine_options(**kw) -> L.KqRefineOptions:
    o = L.KqRefineOptions()
    L.lib().kq_default_refine_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def fcg_refine(b: np.ndarray, x: np.ndarray, sigma: float, parity: int = 0, opt: L.KqRefineOptions | None = None):
    """Refine x in place for (sigma - D D) x = b; returns (status, KqRefineReport)."""
    opt = opt if opt is not None else refine_options()
    rep = L.KqRefineReport()
    st = L.lib().kq_fcg_refine(parity, sigma, _ptr(b), _ptr(x), C.byref(opt), C.byref(rep))
    return st, rep
