"""Parity of the GPU multishift CG with the oracle CGM (inverter_cgm.cpp:17-87)
and with the reference's own known-answer test (test/test_solver.cpp).

Tolerances:
  * reference method, FP64: the oracle's iteration count within 1 % (at least 2:
    the reductions add in a different order, whose grouping also follows the
    autotuned block size) and solutions within 1e-8 relative (both stop at an
    iterated relative residual of 1e-10).
  * every FP64 / mixed solve: recomputed FP64 true relative residual of every
    shift <= the requested 1e-10.
"""
import numpy as np
import pytest

import oracle
from conftest import rel_err


def _true_rel_res(fat, lng, lat, parity, b, x, sigma):
    Ax = sigma * x - oracle.mdaggm(fat, lng, x, lat, parity)
    return float(np.linalg.norm(b - Ax) / np.linalg.norm(b))


@pytest.mark.gpu
def test_reference_known_answer(gpu):
    """test/test_solver.cpp:28-130: 4x4x4x1, FP32, every real = 1, one shift,
    ODD_ODD, tol 1e-4 -> every output real is 0.007561 (mean |diff| < tol)."""
    kq = gpu
    lat = (4, 4, 4, 1)
    V = 64
    Vh = V // 2
    gauge = np.ones((V, 4, 9, 2), dtype=np.float32)
    for lt in (kq.KUQCD_FAT_LINK, kq.KUQCD_LONG_LINK):
        p = kq.gauge_param(lat, prec=kq.KUQCD_SINGLE, link_type=lt, order=kq.KUQCD_NORMAL_ORDER, halo_depth=3,
                           site_subset=kq.KUQCD_ALL, od=kq.KUQCD_TZYX_ORDER)
        assert kq.kuqcd_load_gaugefield(gauge, p) == kq.KUQCD_SUCCESS
    src = np.ones((Vh, 3, 2), dtype=np.float32)
    x = [np.zeros((Vh, 3, 2), dtype=np.float32)]
    prm = kq.InvertParam(shifts=np.array([3.671676313381e-05]), tol=1e-4, max_iter=6500, prec=kq.KUQCD_SINGLE,
                         matpc_type=kq.KUQCD_ODD_ODD)
    assert kq.kuqcd_solve_multishift(x, src, prm) == kq.KUQCD_SUCCESS
    diff = np.abs(x[0] - 0.007561).sum() / (Vh * 6)
    assert diff < 1e-4
    # closed form (1+i)/(132.25 + sigma): see tests/test_oracle.py
    assert np.allclose(x[0], 1.0 / (132.25 + 3.671676313381e-05), rtol=1e-5)
    # the oracle's float CGM gives the same iterate
    xo, it = oracle.cgm(gauge, gauge, lat, 1, np.array([3.671676313381e-05], np.float32), src, 1e-4, 6500)
    assert prm.iter == it
    assert rel_err(x[0], xo[0]) <= 1e-6


def _config1(kq, lat=(8, 8, 8, 8)):
    fat, lng = kq.gen_gauge(lat, kind=kq.KQ_GEN_HAAR, fat_scale=0.9, seed_fat=1, seed_long=2)
    b = kq.gen_source(lat, 0, seed=3)
    m = 0.05
    s = np.array([0.0, 0.01, 0.1, 1.0])
    shifts = np.concatenate([[m * m + s[0]], s[1:] - s[0]])  # reference convention
    sig_abs = m * m + s
    tol = 1e-20 * float(np.sum(b * b))
    return fat, lng, b, shifts, sig_abs, tol


@pytest.mark.gpu
def test_cgm_fp64_matches_oracle_config1(gpu):
    kq = gpu
    lat = (8, 8, 8, 8)
    fat, lng, b, shifts, sig_abs, tol = _config1(kq, lat)
    kq.load_gauge_pair(fat, lng, lat)
    x = [np.zeros_like(b) for _ in shifts]
    prm = kq.InvertParam(shifts=shifts, tol=tol, max_iter=5000, prec=kq.KUQCD_DOUBLE, matpc_type=kq.KUQCD_EVEN_EVEN)
    assert kq.kuqcd_solve_multishift(x, b, prm) == kq.KUQCD_SUCCESS, kq.last_error()
    xo, it = oracle.cgm(fat, lng, lat, 0, shifts, b, tol, 5000)
    print("iterations gpu", prm.iter, "oracle", it)
    assert abs(prm.iter - it) <= max(2, it // 100)
    for k in range(len(shifts)):
        assert rel_err(x[k], xo[k]) <= 1e-8, k
        assert _true_rel_res(fat, lng, lat, 0, b, x[k], sig_abs[k]) <= 1.5e-10


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [1, 2, 3, 4])
def test_cgm_modes_reach_tolerance(gpu, mode):
    kq = gpu
    lat = (8, 8, 8, 8)
    fat, lng, b, shifts, sig_abs, tol = _config1(kq, lat)
    kq.load_gauge_pair(fat, lng, lat)
    x = [np.zeros_like(b) for _ in shifts]
    prm = kq.InvertParam(shifts=shifts, tol=tol, max_iter=5000)
    opt = kq.default_options(mode=mode, refine=1, verbose=0)
    st, rep = kq.solve_ex(x, b, prm, opt)
    assert st == kq.KUQCD_SUCCESS, kq.last_error()
    print("mode", mode, "iter", rep.iter, "reliable", rep.reliable_updates, "refine", list(rep.refine_iter[:4]),
          "res", list(rep.true_rel_res[:4]))
    xo, _ = oracle.cgm(fat, lng, lat, 0, shifts, b, tol, 5000)
    for k in range(len(shifts)):
        r = _true_rel_res(fat, lng, lat, 0, b, x[k], sig_abs[k])
        assert r <= 1e-10, (k, r)
        assert abs(r - rep.true_rel_res[k]) <= 1e-3 * r + 1e-15
        # error against the FP64 oracle solution, per shift
        assert rel_err(x[k], xo[k]) <= 1e-7, k


@pytest.mark.gpu
@pytest.mark.parametrize("mode,nt,nt64", [(2, 256, 64), (4, 256, 64), (3, 128, 64), (4, 64, 256)])
def test_mixed_block_sizes_of_the_two_hopping_terms(gpu, monkeypatch, mode, nt, nt64):
    """The iteration's hopping term and the FP64 one of the reliable updates are tuned
    independently and write their per-system partial sums into one buffer of the runner, so a
    smaller FP64 block size (more partials per system) must still fit every system's row.
    Forced apart here, with four shifts forking into separate systems; every shift still
    reaches the tolerance and agrees with the FP64 oracle."""
    kq = gpu
    lat = (8, 8, 8, 8)
    fat, lng, b, shifts, sig_abs, tol = _config1(kq, lat)
    kq.load_gauge_pair(fat, lng, lat)
    monkeypatch.setenv("KQ_RELIABLE_NT", str(nt64))
    x = [np.zeros_like(b) for _ in shifts]
    prm = kq.InvertParam(shifts=shifts, tol=tol, max_iter=5000)
    print("mode", mode, "nt", nt, "nt64", nt64, flush=True)
    st, rep = kq.solve_ex(x, b, prm, kq.default_options(mode=mode, nt=nt, refine=1, verbose=0))
    assert st == kq.KUQCD_SUCCESS, kq.last_error()
    assert rep.reliable_updates > 0
    xo, _ = oracle.cgm(fat, lng, lat, 0, shifts, b, tol, 5000)
    for k in range(len(shifts)):
        assert _true_rel_res(fat, lng, lat, 0, b, x[k], sig_abs[k]) <= 1e-10, k
        assert rel_err(x[k], xo[k]) <= 1e-7, k
    assert kq.kq_synchronize() == 0


@pytest.mark.gpu
def test_repeated_solves_reuse_workspace(gpu):
    """Solves of different shapes in a row (cached device buffers are reused by size, the
    autotuner runs again after each gauge reload): every solve reaches the tolerance and agrees
    with the first solve of its shape (the autotuner may pick another block size, which
    regroups the reduction partial sums, so agreement is to rounding, not bitwise)."""
    kq = gpu
    results = {}
    for rnd in range(2):
        for lat in [(8, 8, 8, 8), (4, 4, 4, 8), (8, 8, 8, 8)]:
            fat, lng, b, shifts, sig_abs, tol = _config1(kq, lat)
            kq.load_gauge_pair(fat, lng, lat)
            for mode in (1, 4):
                x = [np.zeros_like(b) for _ in shifts]
                prm = kq.InvertParam(shifts=shifts, tol=tol, max_iter=5000)
                st, rep = kq.solve_ex(x, b, prm, kq.default_options(mode=mode, refine=1, verbose=0))
                assert st == kq.KUQCD_SUCCESS, kq.last_error()
                assert max(rep.true_rel_res[k] for k in range(len(shifts))) <= 1e-10
                key = (lat, mode)
                if key in results:
                    assert max(rel_err(a, c) for a, c in zip(x, results[key])) <= 1e-9
                else:
                    results[key] = x
            kq.kuqcd_gauge_free()  # also releases the cached device memory
    assert kq.kq_synchronize() == 0


@pytest.mark.gpu
def test_cgm_batched_matches_single(gpu):
    kq = gpu
    lat = (8, 4, 4, 8)
    fat, lng = kq.gen_gauge(lat, kind=kq.KQ_GEN_WEAK, eps=0.3)
    kq.load_gauge_pair(fat, lng, lat)
    rng = np.random.default_rng(5)
    nrhs, ns = 4, 6
    bs = [kq.gen_source(lat, 0, seed=1000 + r) for r in range(nrhs)]
    sh = np.zeros((nrhs, ns))
    for r in range(nrhs):
        s = np.sort(10 ** rng.uniform(-4, 0, ns))
        sh[r] = np.concatenate([[s[0]], s[1:] - s[0]])
    xs = [np.zeros_like(bs[0]) for _ in range(nrhs * ns)]
    prm = kq.InvertParam(shifts=sh, tol=0.0, max_iter=8000)
    opt = kq.default_options(mode=2, refine=1, rel_tol=1e-10, verbose=0)
    st, reps = kq.solve_batched(xs, bs, prm, opt)
    assert st == kq.KUQCD_SUCCESS, kq.last_error()
    for r in range(nrhs):
        sig_abs = np.concatenate([[sh[r, 0]], sh[r, 0] + sh[r, 1:]])
        for k in range(ns):
            res = _true_rel_res(fat, lng, lat, 0, bs[r], xs[r * ns + k], sig_abs[k])
            assert res <= 1e-10, (r, k, res)


@pytest.mark.gpu
def test_status_codes_gpu(gpu):
    kq = gpu
    lat = (4, 4, 4, 4)
    b = kq.gen_source(lat, 0)
    x = [np.zeros_like(b)]
    prm = kq.InvertParam(shifts=np.array([0.1]), tol=1e-12)
    # solve before any gauge is loaded
    assert kq.kuqcd_solve_multishift(x, b, prm) == kq.KUQCD_ERR_UNKNOWN_TYPE
    fat, lng = kq.gen_gauge(lat)
    kq.load_gauge_pair(fat, lng, lat)
    for field, val, want in [("inverter_type", kq.KUQCD_CG, kq.KUQCD_ERR_UNSUPPORTED),
                             ("dslash_type", kq.KUQCD_WILSON_DIRAC, kq.KUQCD_ERR_UNSUPPORTED),
                             ("dslash_type", kq.KUQCD_INVALID_DIRAC, kq.KUQCD_ERR_INVALID_PARAM),
                             ("matpc_type", 7, kq.KUQCD_ERR_INVALID_PARAM),
                             ("prec", kq.KUQCD_HALF, kq.KUQCD_ERR_UNSUPPORTED),
                             ("prec", 3, kq.KUQCD_ERR_INVALID_PARAM),
                             ("prec", kq.KUQCD_SINGLE, kq.KUQCD_ERR_INVALID_PARAM)]:
        p = kq.InvertParam(shifts=np.array([0.1]), tol=1e-12)
        setattr(p, field, val)
        assert kq.kuqcd_solve_multishift(x, b, p) == want, (field, val)
    p = kq.InvertParam(shifts=np.array([0.1]), tol=1e-12, num_shift=0)
    assert kq.kuqcd_solve_multishift(x, b, p) == kq.KUQCD_ERR_INVALID_PARAM
    # max_iter exhaustion is not an error (inverter_cgm.cpp:80-81)
    p = kq.InvertParam(shifts=np.array([0.001]), tol=1e-30, max_iter=3)
    assert kq.kuqcd_solve_multishift(x, b, p) == kq.KUQCD_SUCCESS
    assert p.iter == 3
    # max_iter 0: the reference's do-while still runs one iteration
    p = kq.InvertParam(shifts=np.array([0.001]), tol=1e-30, max_iter=0)
    assert kq.kuqcd_solve_multishift(x, b, p) == kq.KUQCD_SUCCESS
    assert p.iter == 1


@pytest.mark.gpu
def test_cpp_dropin_caller():
    """An unmodified C++ caller of include/kuqcd.h (tests/cpp/dropin_test_solver.cpp)
    runs the reference's known-answer case, reading the grid from stdin as
    test/run.sh feeds test/input."""
    import subprocess
    from pathlib import Path
    here = Path(__file__).resolve().parent / "cpp"
    subprocess.run(["make", "-s", "-C", str(here)], check=True)
    r = subprocess.run([str(here / "dropin_test_solver")], input="gridx=1\ngridy=1\ngridz=1\ngridt=1\n",
                       capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASSED" in r.stdout
