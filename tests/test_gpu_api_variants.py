"""The drop-in API's accepted input variants, through kuqcd_solve_multishift:

  * BATCH_ORDER gauges (kuqcd.cpp:343-368): the host layout is the NORMAL one
    (kuqcd.cpp:118-121, the Nbatch lanes only change the reference's internal
    storage), so a pair loaded in BATCH order solves to the same solutions as the
    same pair loaded in NORMAL order, in float (Nbatch 16) and double (Nbatch 8).
    The lattices have Lx divisible by Nbatch, as latt_base.cpp:77-93 asks.
  * dispatch order when more than one gauge pair is loaded (kuqcd.cpp:415-435):
    float/BATCH, then double/BATCH, then float/NORMAL, then double/NORMAL. Each slot
    holds a different gauge here, so the solution shows which one was used.
  * halo_depth 0 (kuqcd.cpp:62-99): at one rank no halo exists, so depth 0 and 3
    give the same operator.
  * the float reference method (KUQCD_SINGLE, Q4: float recurrence scalars) over
    hundreds of iterations against the oracle's float CGM (inverter_cgm.cpp:17-87).
  * KUQCD_HISQ_DIRAC (the non-PC operator), which the reference runs on fields of one parity
    that its operator indexes as fields of both (out of bounds, dirac_hisq.cpp:130, 422): here
    b and x_k hold both parities (V sites, even block first) and CGM::Invert runs with the full
    operator; against the oracle's CGM on fields of both parities (pinned by dense solves in
    tests/test_oracle.py) in double and float, and each parity of the solution against the
    library's own even-odd solve of that parity. The extensions reject it.

Tolerances: a pair in BATCH and NORMAL order, and halo depth 0 and 3, run the
same kernels with the same fixed reduction order, so their solutions are
compared bitwise. The float method is compared with the oracle's float CGM,
which adds its dot products in float while the library adds them in FP64: the
iteration counts within 1 % (at least 2), and the two float solutions within
5e-6 of each other, where each lies ~7e-7 from the FP64 solution at this
tolerance (measured with the oracle in both precisions).
"""
import numpy as np
import pytest

import oracle
from conftest import rel_err


def _problem(kq, lat, seed, prec):
    fat, lng = kq.gen_gauge(lat, kind=kq.KQ_GEN_HAAR, fat_scale=0.9, seed_fat=seed, seed_long=seed + 1)
    b = kq.gen_source(lat, 0, seed=seed + 2)
    m = 0.05
    s = np.array([0.0, 0.01, 0.1, 1.0])
    shifts = np.concatenate([[m * m + s[0]], s[1:] - s[0]])
    dt = np.float64 if prec == kq.KUQCD_DOUBLE else np.float32
    return fat.astype(dt), lng.astype(dt), b.astype(dt), shifts


def _solve(kq, b, shifts, rel, prec, max_iter=5000):
    tol = rel * rel * float(np.sum(b.astype(np.float64) ** 2))
    x = [np.zeros_like(b) for _ in shifts]
    prm = kq.InvertParam(shifts=shifts, tol=tol, max_iter=max_iter, prec=prec, matpc_type=kq.KUQCD_EVEN_EVEN)
    st = kq.kuqcd_solve_multishift(x, b, prm)
    return st, x, prm.iter, tol


def _load(kq, fat, lng, lat, prec, order, halo_depth=3):
    for arr, lt in ((fat, kq.KUQCD_FAT_LINK), (lng, kq.KUQCD_LONG_LINK)):
        p = kq.gauge_param(lat, prec=prec, link_type=lt, order=order, halo_depth=halo_depth)
        assert kq.kuqcd_load_gaugefield(np.ascontiguousarray(arr), p) == kq.KUQCD_SUCCESS, kq.last_error()


@pytest.mark.gpu
@pytest.mark.parametrize("prec,lat", [(4, (16, 4, 4, 8)), (8, (8, 8, 8, 8))])
def test_batch_order_solves_like_normal(gpu, prec, lat):
    kq = gpu
    fat, lng, b, shifts = _problem(kq, lat, 11, prec)
    rel = 1e-5 if prec == kq.KUQCD_SINGLE else 1e-10
    out = {}
    for order in (kq.KUQCD_NORMAL_ORDER, kq.KUQCD_BATCH_ORDER):
        print("order", order, "prec", prec, "lattice", lat, flush=True)
        _load(kq, fat, lng, lat, prec, order)
        st, x, it, _ = _solve(kq, b, shifts, rel, prec)
        assert st == kq.KUQCD_SUCCESS, kq.last_error()
        out[order] = (x, it)
        kq.kuqcd_gauge_free()
    (xn, itn), (xb, itb) = out[kq.KUQCD_NORMAL_ORDER], out[kq.KUQCD_BATCH_ORDER]
    assert itn == itb
    for k in range(len(shifts)):
        assert np.array_equal(xn[k], xb[k]), k
    # and both are the oracle's solution
    xo, ito = oracle.cgm(fat, lng, lat, 0, shifts.astype(b.dtype), b, rel * rel * float(np.sum(b.astype(np.float64) ** 2)),
                         5000)
    assert abs(itn - ito) <= max(2, ito // 100), (itn, ito)
    for k in range(len(shifts)):
        assert rel_err(xn[k], xo[k]) <= (5e-6 if prec == kq.KUQCD_SINGLE else 1e-8), k


@pytest.mark.gpu
def test_dispatch_order_of_loaded_pairs(gpu):
    """kuqcd.cpp:415-435: the first loaded pair in the order f16 (float BATCH), d8 (double
    BATCH), f1 (float NORMAL), d1 (double NORMAL) is solved with. The solve's real type is
    that pair's, so a prec that does not match it is rejected."""
    kq = gpu
    lat = (16, 8, 8, 8)
    F, D = kq.KUQCD_SINGLE, kq.KUQCD_DOUBLE
    # one gauge per slot (seeds 21, 31, 41, 51), the same source seed
    probs = {("d1"): _problem(kq, lat, 21, D), ("f1"): _problem(kq, lat, 31, F),
             ("d8"): _problem(kq, lat, 41, D), ("f16"): _problem(kq, lat, 51, F)}
    order_of = {"d1": kq.KUQCD_NORMAL_ORDER, "f1": kq.KUQCD_NORMAL_ORDER, "d8": kq.KUQCD_BATCH_ORDER,
                "f16": kq.KUQCD_BATCH_ORDER}
    prec_of = {"d1": D, "f1": F, "d8": D, "f16": F}
    rel_of = {D: 1e-10, F: 1e-5}

    def solve_with(prec, key_for_b):
        fat, lng, b, shifts = probs[key_for_b]
        return _solve(kq, b, shifts, rel_of[prec], prec)

    # each pair alone: its own solution
    alone = {}
    for key in ("d1", "f1", "d8", "f16"):
        fat, lng, b, shifts = probs[key]
        _load(kq, fat, lng, lat, prec_of[key], order_of[key])
        st, x, it, _ = solve_with(prec_of[key], key)
        assert st == kq.KUQCD_SUCCESS, (key, kq.last_error())
        alone[key] = (x, it)
        kq.kuqcd_gauge_free()

    # load in the reverse of the dispatch order; after each load the newest pair wins
    for n, key in enumerate(("d1", "f1", "d8", "f16")):
        fat, lng, b, shifts = probs[key]
        _load(kq, fat, lng, lat, prec_of[key], order_of[key])
        print("loaded", key, flush=True)
        st, x, it, _ = solve_with(prec_of[key], key)
        assert st == kq.KUQCD_SUCCESS, (key, kq.last_error())
        assert it == alone[key][1], key
        for k in range(len(shifts)):
            assert np.array_equal(x[k], alone[key][0][k]), (key, k)
        # the other real type is rejected once a pair of this type takes precedence
        other = F if prec_of[key] == D else D
        other_key = "f1" if other == F else "d1"
        st, _, _, _ = solve_with(other, other_key)
        assert st == kq.KUQCD_ERR_INVALID_PARAM, (key, st)
    kq.kuqcd_gauge_free()
    # a pair is used only when both its fat and long links are loaded (kuqcd.cpp:415)
    fat, lng, b, shifts = probs["d1"]
    _load(kq, fat, lng, lat, D, kq.KUQCD_NORMAL_ORDER)
    ff, _, _, _ = probs["f16"]
    p = kq.gauge_param(lat, prec=F, link_type=kq.KUQCD_FAT_LINK, order=kq.KUQCD_BATCH_ORDER)
    assert kq.kuqcd_load_gaugefield(np.ascontiguousarray(ff), p) == kq.KUQCD_SUCCESS
    st, x, it, _ = solve_with(D, "d1")
    assert st == kq.KUQCD_SUCCESS and it == alone["d1"][1]
    for k in range(len(shifts)):
        assert np.array_equal(x[k], alone["d1"][0][k]), k


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [4, 8])
def test_halo_depth_zero(gpu, prec):
    kq = gpu
    lat = (8, 8, 8, 8)
    fat, lng, b, shifts = _problem(kq, lat, 61, prec)
    rel = 1e-5 if prec == kq.KUQCD_SINGLE else 1e-10
    out = []
    for depth in (3, 0):
        _load(kq, fat, lng, lat, prec, kq.KUQCD_NORMAL_ORDER, halo_depth=depth)
        st, x, it, _ = _solve(kq, b, shifts, rel, prec)
        assert st == kq.KUQCD_SUCCESS, kq.last_error()
        out.append((x, it))
        kq.kuqcd_gauge_free()
    assert out[0][1] == out[1][1]
    for k in range(len(shifts)):
        assert np.array_equal(out[0][0][k], out[1][0][k]), k


@pytest.mark.gpu
@pytest.mark.parametrize("rel", [1e-4, 1e-5])
def test_float_reference_method_matches_oracle_float_cgm(gpu, rel):
    """The 8^4 problem of the CPU-runnable workload in float (233 and 287 iterations)."""
    kq = gpu
    lat = (8, 8, 8, 8)
    F = kq.KUQCD_SINGLE
    fat, lng, b, shifts = _problem(kq, lat, 1, F)
    _load(kq, fat, lng, lat, F, kq.KUQCD_NORMAL_ORDER)
    st, x, it, tol = _solve(kq, b, shifts, rel, F)
    assert st == kq.KUQCD_SUCCESS, kq.last_error()
    xo, ito = oracle.cgm(fat, lng, lat, 0, shifts.astype(np.float32), b, tol, 5000)
    fd, ld, bd = fat.astype(np.float64), lng.astype(np.float64), b.astype(np.float64)
    xd, itd = oracle.cgm(fd, ld, lat, 0, shifts, bd, tol, 5000)
    print(f"rel {rel}: iterations gpu {it}, oracle float {ito}, oracle double {itd}", flush=True)
    assert it > 100
    assert abs(it - ito) <= max(2, ito // 100), (it, ito)
    for k in range(len(shifts)):
        e_o, e_d = rel_err(x[k], xo[k]), rel_err(x[k], xd[k])
        print(f"shift {k}: vs oracle float {e_o:.2e}, vs oracle double {e_d:.2e}, oracle float vs double "
              f"{rel_err(xo[k], xd[k]):.2e}", flush=True)
        assert e_o <= 5e-6 and e_d <= 5e-6, k


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [8, 4])
def test_full_operator_solve(gpu, prec):
    """The 8^4 problem of the CPU-runnable workload with a source on both parities."""
    kq = gpu
    lat = (8, 8, 8, 8)
    fat, lng, _, shifts = _problem(kq, lat, 71, prec)
    dt = fat.dtype
    b = np.concatenate([kq.gen_source(lat, 0, seed=73), kq.gen_source(lat, 1, seed=74)]).astype(dt)
    Vh = b.shape[0] // 2
    _load(kq, fat, lng, lat, prec, kq.KUQCD_NORMAL_ORDER)
    rel = 1e-5 if prec == kq.KUQCD_SINGLE else 1e-10
    tol = rel * rel * float(np.sum(b.astype(np.float64) ** 2))
    x = [np.zeros_like(b) for _ in shifts]
    prm = kq.InvertParam(shifts=shifts, tol=tol, max_iter=5000, prec=prec, dslash_type=kq.KUQCD_HISQ_DIRAC)
    print("full operator, prec", prec, flush=True)
    assert kq.kuqcd_solve_multishift(x, b, prm) == kq.KUQCD_SUCCESS, kq.last_error()
    xo, ito = oracle.cgm(fat, lng, lat, -1, shifts.astype(dt), b, tol, 5000)
    print(f"iterations gpu {prm.iter}, oracle {ito}", flush=True)
    assert prm.iter > 50 and abs(prm.iter - ito) <= max(2, ito // 100), (prm.iter, ito)
    for k in range(len(shifts)):
        assert rel_err(x[k], xo[k]) <= (5e-6 if prec == kq.KUQCD_SINGLE else 1e-8), k
    # each parity of the solution solves that parity's even-odd system
    for q, mt in ((0, kq.KUQCD_EVEN_EVEN), (1, kq.KUQCD_ODD_ODD)):
        bq = np.ascontiguousarray(b[q * Vh:(q + 1) * Vh])
        rq = rel if prec == kq.KUQCD_SINGLE else rel / 10
        tq = rq * rq * float(np.sum(bq.astype(np.float64) ** 2))
        xq = [np.zeros_like(bq) for _ in shifts]
        pq = kq.InvertParam(shifts=shifts, tol=tq, max_iter=5000, prec=prec, matpc_type=mt)
        assert kq.kuqcd_solve_multishift(xq, bq, pq) == kq.KUQCD_SUCCESS, kq.last_error()
        for k in range(len(shifts)):
            assert rel_err(x[k][q * Vh:(q + 1) * Vh], xq[k]) <= (2e-5 if prec == kq.KUQCD_SINGLE else 1e-8), (q, k)
    # the extended entry points solve the even-odd operator only
    if prec == kq.KUQCD_DOUBLE:
        st, _ = kq.solve_ex(x, b, prm, kq.default_options())
        assert st == kq.KUQCD_ERR_UNSUPPORTED
    kq.kuqcd_gauge_free()
