"""kq_mdaggm_noalloc (no counterpart in the reference; for callers that cannot allocate device
memory) against kq_mdaggm, the entry point whose kernels it shares: two calls in a row, the
second on a smaller lattice, then a float gauge; the results are bitwise equal (same kernels,
block size and reduction order). Also the rejections: a lattice whose scratch exceeds the fixed
64 MiB, host pointers, overlapping src and dst."""
import ctypes as C

import numpy as np
import pytest


def _unit_links(lat):
    V = int(np.prod(lat))
    g = np.zeros((V, 4, 9, 2))
    g[:, :, [0, 4, 8], 0] = 1.0
    return g


def _run_noalloc(kq, src, parity, sigma):
    L = kq.L.lib()
    n = src.nbytes
    d_src, d_dst = L.kq_device_alloc(n), L.kq_device_alloc(n)
    assert d_src and d_dst
    try:
        assert L.kq_memcpy(d_src, src.ctypes.data, n) == 0
        st = L.kq_mdaggm_noalloc(C.c_void_p(d_src), C.c_void_p(d_dst), parity, sigma)
        assert st == kq.KUQCD_SUCCESS, kq.last_error()
        assert L.kq_synchronize() == 0
        out = np.empty_like(src)
        assert L.kq_memcpy(out.ctypes.data, d_dst, n) == 0
        return out
    finally:
        L.kq_device_free(d_src)
        L.kq_device_free(d_dst)


@pytest.mark.gpu
def test_mdaggm_noalloc_matches_mdaggm(gpu):
    kq = gpu
    cases = [((16, 16, 16, 16), kq.KUQCD_DOUBLE, 0, 0.0123), ((8, 8, 8, 8), kq.KUQCD_DOUBLE, 1, 0.5),
             ((8, 4, 4, 8), kq.KUQCD_SINGLE, 0, 0.25)]
    for lat, prec, parity, sigma in cases:
        print("no-alloc operator:", lat, "prec", prec, "parity", parity, flush=True)
        fat, lng = kq.gen_gauge(lat, kind=kq.KQ_GEN_WEAK, eps=0.3, fat_scale=0.9, seed_fat=7, seed_long=8)
        kq.kuqcd_gauge_free()
        kq.load_gauge_pair(fat, lng, lat, prec=prec)
        dt = np.float64 if prec == kq.KUQCD_DOUBLE else np.float32
        src = kq.gen_source(lat, parity, seed=9).astype(dt)
        ref = np.empty_like(src)
        assert kq.mdaggm(src, ref, parity, sigma, kq.KQ_MODE_REFERENCE) == kq.KUQCD_SUCCESS, kq.last_error()
        out = _run_noalloc(kq, src, parity, sigma)
        assert np.array_equal(out, ref), float(np.abs(out - ref).max())


@pytest.mark.gpu
def test_mdaggm_noalloc_rejections(gpu):
    kq = gpu
    L = kq.L.lib()
    lat = (8, 8, 8, 8)
    fat, lng = kq.gen_gauge(lat, kind=kq.KQ_GEN_WEAK, eps=0.3)
    kq.load_gauge_pair(fat, lng, lat)
    src = kq.gen_source(lat, 0, seed=1)
    dst = np.empty_like(src)
    # host memory
    assert L.kq_mdaggm_noalloc(C.c_void_p(src.ctypes.data), C.c_void_p(dst.ctypes.data), 0, 0.1) == \
        kq.KUQCD_ERR_INVALID_PARAM
    d = L.kq_device_alloc(2 * src.nbytes)
    try:
        # overlapping src and dst
        assert L.kq_mdaggm_noalloc(C.c_void_p(d), C.c_void_p(d + 8), 0, 0.1) == kq.KUQCD_ERR_INVALID_PARAM
        assert L.kq_mdaggm_noalloc(None, C.c_void_p(d), 0, 0.1) == kq.KUQCD_ERR_NULL_POINTER
        assert L.kq_mdaggm_noalloc(C.c_void_p(d), C.c_void_p(d + src.nbytes), 2, 0.1) == kq.KUQCD_ERR_INVALID_PARAM
    finally:
        L.kq_device_free(d)
    # 36^4 in double needs 80.6 MB of scratch: above the fixed 64 MiB
    big = (36, 36, 36, 36)
    u = _unit_links(big)
    kq.kuqcd_gauge_free()
    kq.load_gauge_pair(u, u, big)
    n = int(np.prod(big)) // 2 * 6 * 8
    a, b = L.kq_device_alloc(n), L.kq_device_alloc(n)
    try:
        assert L.kq_mdaggm_noalloc(C.c_void_p(a), C.c_void_p(b), 0, 0.1) == kq.KUQCD_ERR_INVALID_PARAM
        assert "64 MiB" in kq.last_error()
    finally:
        L.kq_device_free(a)
        L.kq_device_free(b)
