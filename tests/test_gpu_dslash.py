"""Parity of the HIP hopping term with the oracle (dirac_hisq.cpp:409-502).

Tolerances (relative 2-norm over the whole output vector):
  FP64 links/vectors            <= 1e-13
  FP32 links/vectors            <= 1e-5
  FP16 links (+FP32 scale), FP32 vectors  <= 3e-3 (half precision has an 11-bit significand)
"""
import numpy as np
import pytest

import oracle
from conftest import rel_err

LATS = [(4, 4, 4, 4), (8, 8, 8, 8), (6, 4, 2, 8), (4, 4, 4, 1), (2, 2, 2, 2), (4, 6, 1, 3)]


def _setup(kq, lat, prec, kind=0, seed=11):
    fat, lng = kq.gen_gauge(lat, kind=kind, seed_fat=seed, seed_long=seed + 1)
    if prec == kq.KUQCD_SINGLE:
        fat = fat.astype(np.float32)
        lng = lng.astype(np.float32)
    kq.load_gauge_pair(fat, lng, lat, prec=prec)
    return fat, lng


@pytest.mark.gpu
@pytest.mark.parametrize("lat", LATS)
@pytest.mark.parametrize("parity", [0, 1])
def test_dslash_fp64(gpu, lat, parity):
    print("case", lat, parity, flush=True)
    fat, lng = _setup(gpu, lat, gpu.KUQCD_DOUBLE)
    src = gpu.gen_source(lat, 1 - parity, seed=5)
    dst = np.empty_like(src)
    assert gpu.dslash(src, dst, parity) == 0, gpu.last_error()
    ref = oracle.dslash(fat, lng, src, lat, parity)
    assert rel_err(dst, ref) <= 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("lat", [(4, 4, 4, 4), (8, 8, 8, 8), (4, 4, 4, 1)])
def test_dslash_fp32_gauge(gpu, lat):
    fat, lng = _setup(gpu, lat, gpu.KUQCD_SINGLE)
    src = gpu.gen_source(lat, 1, seed=6).astype(np.float32)
    dst = np.empty_like(src)
    assert gpu.dslash(src, dst, 0) == 0, gpu.last_error()
    ref32 = oracle.dslash(fat, lng, src, lat, 0)
    ref64 = oracle.dslash(fat.astype(np.float64), lng.astype(np.float64), src.astype(np.float64), lat, 0)
    assert rel_err(dst, ref32) <= 1e-5
    assert rel_err(dst, ref64) <= 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("mode,tol", [(2, 1e-5), (3, 3e-3), (4, 5e-4)])
@pytest.mark.parametrize("kind", [0, 1])
def test_dslash_lower_precision_modes(gpu, mode, tol, kind):
    lat = (8, 8, 8, 8)
    fat, lng = _setup(gpu, lat, gpu.KUQCD_DOUBLE, kind=kind)
    src = gpu.gen_source(lat, 0, seed=7)
    dst = np.empty_like(src)
    assert gpu.dslash(src, dst, 1, mode) == 0, gpu.last_error()
    ref = oracle.dslash(fat, lng, src, lat, 1)
    assert rel_err(dst, ref) <= tol


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [1, 2, 3, 4])
def test_dslash_batched_equals_single(gpu, mode):
    lat = (8, 4, 4, 8)
    _setup(gpu, lat, gpu.KUQCD_DOUBLE)
    nrhs = 5
    src = np.stack([gpu.gen_source(lat, 1, seed=100 + r) for r in range(nrhs)])
    single = np.empty_like(src)
    for r in range(nrhs):
        assert gpu.dslash(np.ascontiguousarray(src[r]), single[r], 0, mode) == 0
    nrs = [1, 2] if mode == 1 else [1, 2, 3, 4, 8]
    for nr in nrs:
        for nt in (64, 128, 256):
            out = np.empty_like(src)
            assert gpu.dslash_batched(src, out, 0, mode, nr, nt) == 0, gpu.last_error()
            # same arithmetic per right-hand side; the compiler may contract
            # a*b+c into FMAs differently per instantiation, so allow rounding
            assert rel_err(out, single) <= (1e-15 if mode == 1 else 1e-6), (nr, nt)


@pytest.mark.gpu
@pytest.mark.parametrize("lat", [(32, 16, 8, 8), (16, 32, 8, 8), (32, 32, 4, 8)])
@pytest.mark.parametrize("mode", [1, 2, 3, 4])
def test_dslash_batched_staged_rows(gpu, lat, mode):
    """Lattices whose 256-thread workgroups cover whole x-rows of one (x, y) plane, where the
    batched terms (NR >= 2) stage the rows and their +-3 y-neighbour rows in LDS: one workgroup
    per plane with more staged rows than the plane has (the y wrap repeats rows), and two
    workgroups per plane. Same results as single right-hand sides (bitwise the same arithmetic)
    and as the oracle."""
    print("case", lat, mode, flush=True)
    fat, lng = _setup(gpu, lat, gpu.KUQCD_DOUBLE)
    nrhs = 5
    src = np.stack([gpu.gen_source(lat, 0, seed=200 + r) for r in range(nrhs)])
    single = np.empty_like(src)
    for r in range(nrhs):
        assert gpu.dslash(np.ascontiguousarray(src[r]), single[r], 1, mode) == 0
    tol = {1: 1e-13, 2: 1e-5, 3: 3e-3, 4: 5e-4}[mode]
    for r in (0, nrhs - 1):
        assert rel_err(single[r], oracle.dslash(fat, lng, src[r], lat, 1)) <= tol
    for nr in ([2] if mode == 1 else [2, 3, 4, 8]):
        out = np.empty_like(src)
        assert gpu.dslash_batched(src, out, 1, mode, nr, 256) == 0, gpu.last_error()
        assert rel_err(out, single) <= (1e-15 if mode == 1 else 1e-6), nr


_UNSTAGED_CHILD = """
import sys
import numpy as np
sys.path.insert(0, {root!r})
import kuqcd_amd.api as A
A.L.load()
A.set_verbosity(0)
assert A.kuqcd_init() == A.KUQCD_SUCCESS
lat = {lat!r}
fat, lng = A.gen_gauge(lat, kind=0, seed_fat=11, seed_long=12)
A.load_gauge_pair(fat, lng, lat, prec=A.KUQCD_DOUBLE)
src = np.stack([A.gen_source(lat, 0, seed=300 + r) for r in range({nrhs})])
out = np.empty_like(src)
assert A.dslash_batched(src, out, 1, {mode}, {nr}, 256) == 0, A.last_error()
np.save({path!r}, out)
A.kuqcd_gauge_free()
A.kuqcd_end()
"""


@pytest.mark.gpu
@pytest.mark.parametrize("mode,nr", [(1, 2), (2, 4), (2, 8), (4, 3)])
def test_dslash_staged_rows_bitwise_equal_to_unstaged(gpu, tmp_path, mode, nr):
    """The LDS-staged batched term (and its link prefetch for 4 or more FP32 right-hand sides)
    performs the unstaged term's arithmetic in the same order: the outputs are bitwise equal to
    those of a process with the staging switched off (KQ_STAGE=0)."""
    import os
    import subprocess
    import sys
    from conftest import ROOT

    lat, nrhs = (32, 16, 8, 8), 5
    print("case", mode, nr, flush=True)
    _setup(gpu, lat, gpu.KUQCD_DOUBLE)
    src = np.stack([gpu.gen_source(lat, 0, seed=300 + r) for r in range(nrhs)])
    out = np.empty_like(src)
    assert gpu.dslash_batched(src, out, 1, mode, nr, 256) == 0, gpu.last_error()
    path = str(tmp_path / "unstaged.npy")
    code = _UNSTAGED_CHILD.format(root=str(ROOT), lat=lat, nrhs=nrhs, mode=mode, nr=nr, path=path)
    r = subprocess.run([sys.executable, "-c", code], env={**os.environ, "KQ_STAGE": "0"}, capture_output=True,
                       text=True, timeout=180)
    assert r.returncode == 0, r.stderr[-2000:]
    assert np.array_equal(out, np.load(path))


@pytest.mark.gpu
@pytest.mark.parametrize("parity", [0, 1])
def test_mdaggm_fp64(gpu, parity):
    lat = (8, 8, 8, 8)
    fat, lng = _setup(gpu, lat, gpu.KUQCD_DOUBLE)
    src = gpu.gen_source(lat, parity, seed=8)
    dst = np.empty_like(src)
    sigma = 0.0123
    assert gpu.mdaggm(src, dst, parity, sigma) == 0
    ref = sigma * src - oracle.mdaggm(fat, lng, src, lat, parity)
    assert rel_err(dst, ref) <= 1e-13


@pytest.mark.gpu
def test_dslash_antihermitian_full_size(gpu):
    """<u, D_{p,pbar} v> = -<D_{pbar,p} u, v> at the benchmark lattice 32^4 (size-independent property)."""
    lat = (32, 32, 32, 32)
    _setup(gpu, lat, gpu.KUQCD_DOUBLE, kind=1)
    u = gpu.gen_source(lat, 0, seed=21)
    v = gpu.gen_source(lat, 1, seed=22)
    Dv = np.empty_like(u)
    Du = np.empty_like(v)
    assert gpu.dslash(v, Dv, 0) == 0
    assert gpu.dslash(u, Du, 1) == 0
    cu = u[..., 0] + 1j * u[..., 1]
    cv = v[..., 0] + 1j * v[..., 1]
    a = np.vdot(cu, Dv[..., 0] + 1j * Dv[..., 1])
    b = np.vdot(Du[..., 0] + 1j * Du[..., 1], cv)
    assert abs(a + b) <= 1e-12 * abs(a)


@pytest.mark.gpu
@pytest.mark.parametrize("variant,b_site", [("su3", 2144), ("u3", 2144), ("perturbed", 2400)])
def test_dslash_fp64_long_link_formats(gpu, variant, b_site):
    """FP64 long links are kept in 14 reals (two rows of an SU(3) matrix and a complex scale) when
    every link is a complex multiple of an SU(3) matrix to 1e-13: products of SU(3) links times
    c_Naik, and the same times a U(1) phase per link (U(3) links, as HISQ's reunitarised ones).
    Links that are not (a 1e-6 perturbation) keep 18 reals. The operator is the oracle's to 1e-13
    either way; the algorithmic bytes per site (8 x 144 fat + 8 x 112 or 144 long + 2 x 48)
    show which format ran."""
    kq = gpu
    lat = (8, 8, 8, 8)
    fat, lng = kq.gen_gauge(lat, kind=kq.KQ_GEN_HAAR, seed_fat=21, seed_long=22)
    rng = np.random.default_rng(23)
    if variant == "u3":
        ph = np.exp(1j * rng.uniform(0, 2 * np.pi, lng.shape[:2]))[..., None]  # (V, 4, 1)
        z = (lng[..., 0] + 1j * lng[..., 1]) * ph
        lng = np.ascontiguousarray(np.stack([z.real, z.imag], axis=-1))
    elif variant == "perturbed":
        lng = lng + 1e-6 * rng.standard_normal(lng.shape)
    kq.load_gauge_pair(fat, lng, lat)
    st, r = kq.bench_dslash(kq.KQ_MODE_FP64, 1, 1, 256, 1)
    assert st == 0, kq.last_error()
    assert round(r.bytes / (np.prod(lat) // 2)) == b_site, variant
    for parity in (0, 1):
        src = kq.gen_source(lat, 1 - parity, seed=24)
        dst = np.empty_like(src)
        assert kq.dslash(src, dst, parity) == 0, kq.last_error()
        assert rel_err(dst, oracle.dslash(fat, lng, src, lat, parity)) <= 1e-13, (variant, parity)
    kq.kuqcd_gauge_free()


@pytest.mark.gpu
@pytest.mark.parametrize("variant,b_site", [("su3", 1040), ("u3", 1072), ("perturbed", 1200)])
def test_dslash_fp32_long_link_formats(gpu, variant, b_site):
    """A float pair's long links are kept in 12 reals (two rows of an SU(3) matrix) with a real
    scale when every link is a real multiple of an SU(3) matrix to 1e-6, with a complex scale when
    they are complex multiples (U(3) links with a phase in their determinant); perturbed links keep
    18 reals. The operator is the oracle's float operator to 1e-5 each time (8 x 72 fat + 8 x 52,
    56 or 72 long + 2 x 24 bytes per site)."""
    kq = gpu
    lat = (8, 8, 8, 8)
    fat, lng = kq.gen_gauge(lat, kind=kq.KQ_GEN_HAAR, seed_fat=31, seed_long=32)
    rng = np.random.default_rng(33)
    if variant == "u3":
        ph = np.exp(1j * rng.uniform(0, 2 * np.pi, lng.shape[:2]))[..., None]
        z = (lng[..., 0] + 1j * lng[..., 1]) * ph
        lng = np.ascontiguousarray(np.stack([z.real, z.imag], axis=-1))
    elif variant == "perturbed":
        lng = lng + 1e-4 * rng.standard_normal(lng.shape)
    fat32, lng32 = fat.astype(np.float32), lng.astype(np.float32)
    kq.load_gauge_pair(fat32, lng32, lat, prec=kq.KUQCD_SINGLE)
    st, r = kq.bench_dslash(kq.KQ_MODE_REFERENCE, 1, 1, 256, 1)
    assert st == 0, kq.last_error()
    assert round(r.bytes / (np.prod(lat) // 2)) == b_site, variant
    for parity in (0, 1):
        src = kq.gen_source(lat, 1 - parity, seed=34).astype(np.float32)
        dst = np.empty_like(src)
        assert kq.dslash(src, dst, parity) == 0, kq.last_error()
        assert rel_err(dst, oracle.dslash(fat32, lng32, src, lat, parity)) <= 1e-5, (variant, parity)
    kq.kuqcd_gauge_free()


@pytest.mark.gpu
@pytest.mark.parametrize("mode,tol,b_su3,b_u3", [(2, 1e-5, 1040, 1072), (3, 3e-3, 592, 624), (4, 5e-4, 848, 880)])
def test_dslash_mixed_long_link_forms(gpu, mode, tol, b_su3, b_u3):
    """The mixed modes keep long links in 12 reals with a real scale when they are real multiples
    of SU(3) matrices, and with a complex scale when they are complex multiples (U(3) products, as
    HISQ's reunitarised links give: a phase per link here): 8 x 52 / 28 bytes per long link, or
    8 x 56 / 32. Both agree with the FP64 oracle to the mode's tolerance."""
    kq = gpu
    lat = (8, 8, 8, 8)
    fat, lng = kq.gen_gauge(lat, kind=kq.KQ_GEN_HAAR, seed_fat=41, seed_long=42)
    rng = np.random.default_rng(43)
    ph = np.exp(1j * rng.uniform(0, 2 * np.pi, lng.shape[:2]))[..., None]
    z = (lng[..., 0] + 1j * lng[..., 1]) * ph
    lng_u3 = np.ascontiguousarray(np.stack([z.real, z.imag], axis=-1))
    for links, b_site in ((lng, b_su3), (lng_u3, b_u3)):
        kq.load_gauge_pair(fat, links, lat)
        st, r = kq.bench_dslash(mode, 1, 1, 256, 1)
        assert st == 0, kq.last_error()
        assert round(r.bytes / (np.prod(lat) // 2)) == b_site, (mode, b_site)
        src = kq.gen_source(lat, 0, seed=44)
        dst = np.empty_like(src)
        assert kq.dslash(src, dst, 1, mode) == 0, kq.last_error()
        assert rel_err(dst, oracle.dslash(fat, links, src, lat, 1)) <= tol, (mode, b_site)
        kq.kuqcd_gauge_free()


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [2, 4])
def test_mixed_solve_u3_long_links(gpu, mode):
    """A mixed solve whose long links are U(3) products (12 reals with a complex scale) reaches
    every shift's FP64 target and matches the FP64 solve of the same links."""
    kq = gpu
    lat = (8, 8, 8, 8)
    fat, lng = kq.gen_gauge(lat, kind=kq.KQ_GEN_HAAR, seed_fat=51, seed_long=52)
    rng = np.random.default_rng(53)
    ph = np.exp(1j * rng.uniform(0, 2 * np.pi, lng.shape[:2]))[..., None]
    z = (lng[..., 0] + 1j * lng[..., 1]) * ph
    lng = np.ascontiguousarray(np.stack([z.real, z.imag], axis=-1))
    kq.load_gauge_pair(fat, lng, lat)
    b = kq.gen_source(lat, 0, seed=54)
    shifts = np.array([0.0025, 0.01, 0.1, 1.0])
    out = {}
    for m in (kq.KQ_MODE_FP64, mode):
        x = [np.zeros_like(b) for _ in shifts]
        prm = kq.InvertParam(shifts=shifts, tol=0.0, max_iter=5000)
        st, rep = kq.solve_ex(x, b, prm, kq.default_options(mode=m, refine=1, rel_tol=1e-10))
        assert st == kq.KUQCD_SUCCESS, kq.last_error()
        assert max(rep.true_rel_res[:len(shifts)]) <= 1e-10, m
        out[m] = x
    for k in range(len(shifts)):
        assert rel_err(out[mode][k], out[kq.KQ_MODE_FP64][k]) <= 1e-7, k
    kq.kuqcd_gauge_free()
