"""Pins the oracle (oracle/kq_oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible restatement mistake fails it:
  * known answer of the reference's only test (test/test_solver.cpp:28-130)
  * constant links: D reduces to one 3x3 matrix, M = sum_mu [1/2 (U - U^+) + c3000 (N - N^+)]
    (a dropped term, wrong sign, missing dagger or transposed operand changes M)
  * unit links, plane waves: D e^{ipx} = i sum_mu [sin p_mu - (1/24) sin 3p_mu] e^{ipx}
    (a wrong hop length or direction changes the eigenvalue)
  * anti-Hermiticity <u, D_{e,o} v> = -<D_{o,e} u, v> (backward-hop link placement)
  * gauge covariance under U(x) -> g(x) U(x) g(x+mu)^+, N(x) -> g(x) N(x) g(x+3mu)^+
  * an independent dense-matrix construction from full-lattice coordinates
  * CGM against dense solves of every shifted system, with the reference's
    shift convention (inverter_cgm.cpp:52: system k is shifts[0] + shifts[k] - DD), on one
    parity and on fields of both parities with the full operator
"""
import numpy as np
import pytest

import oracle
from conftest import rel_err

C3000 = -1.0 / 48.0


def _cplx(a):
    return a[..., 0] + 1j * a[..., 1]


def _real(z):
    return np.stack([z.real, z.imag], axis=-1)


def _haar(rng, n):
    z = (rng.standard_normal((n, 3, 3)) + 1j * rng.standard_normal((n, 3, 3))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=1, axis2=2)
    q = q * (d / np.abs(d))[:, None, :]
    det = np.linalg.det(q)
    return q / (det ** (1.0 / 3.0))[:, None, None]


def _random_links(rng, lat, scale=1.0):
    V = int(np.prod(lat))
    return _real(scale * _haar(rng, V * 4).reshape(V, 4, 9)).copy()


def _cb(c, L):
    return (c[0] >> 1) + (L[0] // 2) * (c[1] + L[1] * (c[2] + L[2] * c[3]))


def _sites(L, parity):
    for t in range(L[3]):
        for z in range(L[2]):
            for y in range(L[1]):
                for x in range(L[0]):
                    if (x + y + z + t) % 2 == parity:
                        yield np.array([x, y, z, t])


def dense_D(fat, lng, L, dst_parity):
    """Dense hopping matrix built from full-lattice coordinates (independent of
    the oracle's checkerboard decomposition)."""
    V = int(np.prod(L))
    Vh = V // 2
    Uf = _cplx(fat).reshape(V, 4, 3, 3)
    Un = _cplx(lng).reshape(V, 4, 3, 3)
    M = np.zeros((3 * Vh, 3 * Vh), dtype=complex)
    for x in _sites(L, dst_parity):
        i = _cb(x, L)
        gi = dst_parity * Vh + i
        for mu in range(4):
            for h, U, c in ((1, Uf, 0.5), (3, Un, C3000)):
                y = x.copy()
                y[mu] = (y[mu] + h) % L[mu]
                j = _cb(y, L)
                M[3 * i:3 * i + 3, 3 * j:3 * j + 3] += c * U[gi, mu]
                y = x.copy()
                y[mu] = (y[mu] - h) % L[mu]
                j = _cb(y, L)
                gj = (1 - dst_parity) * Vh + j
                M[3 * i:3 * i + 3, 3 * j:3 * j + 3] -= c * U[gj, mu].conj().T
    return M


def test_known_answer_test_solver():
    """test_solver.cpp: 4x4x4x1, all reals 1, sigma = 3.671676313381e-05, ODD_ODD,
    FP32, tol 1e-4: every output real is 0.007561 (mean |diff| < 1e-4)."""
    lat = (4, 4, 4, 1)
    g = np.ones((64, 4, 9, 2), np.float32)
    b = np.ones((32, 3, 2), np.float32)
    x, it = oracle.cgm(g, g, lat, 1, np.array([3.671676313381e-05], np.float32), b, 1e-4, 6500)
    assert np.abs(x[0] - 0.007561).mean() < 1e-4
    # closed form: U v = 6i, U^+ v = 6 per component for entries 1+i, so
    # D v = 4 (1/2 + c3000)(6i - 6) ... = 11.5 i v, D D v = -132.25 v
    assert np.allclose(x[0], 1.0 / (132.25 + 3.671676313381e-05), rtol=1e-6)
    assert it == 1


@pytest.mark.parametrize("lat", [(4, 4, 4, 1), (4, 4, 4, 4), (2, 3, 1, 5)])
def test_constant_field_closed_form(lat):
    rng = np.random.default_rng(1)
    V = int(np.prod(lat))
    Vh = V // 2
    Uf = rng.standard_normal((4, 3, 3)) + 1j * rng.standard_normal((4, 3, 3))
    Un = rng.standard_normal((4, 3, 3)) + 1j * rng.standard_normal((4, 3, 3))
    fat = np.broadcast_to(_real(Uf.reshape(4, 9)), (V, 4, 9, 2)).copy()
    lng = np.broadcast_to(_real(Un.reshape(4, 9)), (V, 4, 9, 2)).copy()
    v = rng.standard_normal(3) + 1j * rng.standard_normal(3)
    src = np.broadcast_to(_real(v), (Vh, 3, 2)).copy()
    M = sum(0.5 * (Uf[m] - Uf[m].conj().T) + C3000 * (Un[m] - Un[m].conj().T) for m in range(4))
    for p in (0, 1):
        out = _cplx(oracle.dslash(fat, lng, src, lat, p))
        assert np.allclose(out, (M @ v)[None, :], rtol=0, atol=1e-13 * np.abs(M @ v).max())
    # all entries 1+i: D v = 11.5 i v and -D D = 132.25
    ones = np.ones((V, 4, 9, 2))
    one = np.ones((Vh, 3, 2))
    Dv = _cplx(oracle.dslash(ones, ones, one, lat, 0))
    assert np.allclose(Dv, 11.5j * (1 + 1j))
    DDv = _cplx(oracle.mdaggm(ones, ones, one, lat, 0))
    assert np.allclose(DDv, -132.25 * (1 + 1j))


@pytest.mark.parametrize("lat", [(8, 4, 6, 10), (4, 4, 4, 4)])
def test_free_field_plane_waves(lat):
    V = int(np.prod(lat))
    Vh = V // 2
    unit = np.zeros((V, 4, 9, 2))
    unit[:, :, [0, 4, 8], 0] = 1.0
    rng = np.random.default_rng(2)
    for _ in range(4):
        n = [int(rng.integers(0, L)) for L in lat]
        p = [2 * np.pi * n[m] / lat[m] for m in range(4)]
        col = rng.standard_normal(3) + 1j * rng.standard_normal(3)

        def wave(parity):
            out = np.zeros((Vh, 3), complex)
            for s in range(Vh):
                c = oracle.full_coords(lat, parity, s)
                out[s] = np.exp(1j * sum(p[m] * c[m] for m in range(4))) * col
            return out

        lam = 1j * sum(np.sin(p[m]) - np.sin(3 * p[m]) / 24.0 for m in range(4))
        for parity in (0, 1):
            src = _real(wave(1 - parity)).copy()
            out = _cplx(oracle.dslash(unit, unit, src, lat, parity))
            assert np.allclose(out, lam * wave(parity), atol=1e-12)


@pytest.mark.parametrize("lat", [(4, 4, 4, 4), (6, 4, 2, 8), (4, 4, 4, 1)])
def test_antihermitian(lat):
    rng = np.random.default_rng(3)
    fat = _random_links(rng, lat, 0.9)
    lng = _random_links(rng, lat, 1 / 24)
    Vh = int(np.prod(lat)) // 2
    u = rng.standard_normal((Vh, 3, 2))
    v = rng.standard_normal((Vh, 3, 2))
    Dv = oracle.dslash(fat, lng, v, lat, 0)
    Du = oracle.dslash(fat, lng, u, lat, 1)
    a = np.vdot(_cplx(u), _cplx(Dv))
    b = np.vdot(_cplx(Du), _cplx(v))
    assert abs(a + b) < 1e-12 * abs(a)


def test_gauge_covariance():
    lat = (4, 6, 4, 4)
    V = int(np.prod(lat))
    Vh = V // 2
    rng = np.random.default_rng(4)
    fat = _random_links(rng, lat, 0.9)
    lng = _random_links(rng, lat, 1 / 24)
    g = _haar(rng, V)  # g at storage index i = q*Vh + s

    def idx(c):
        return ((sum(c) % 2) * Vh) + _cb(c, lat)

    Uf = _cplx(fat).reshape(V, 4, 3, 3)
    Un = _cplx(lng).reshape(V, 4, 3, 3)
    Uf2 = np.empty_like(Uf)
    Un2 = np.empty_like(Un)
    for q in (0, 1):
        for x in _sites(lat, q):
            i = idx(x)
            for mu in range(4):
                y1 = x.copy()
                y1[mu] = (y1[mu] + 1) % lat[mu]
                y3 = x.copy()
                y3[mu] = (y3[mu] + 3) % lat[mu]
                Uf2[i, mu] = g[i] @ Uf[i, mu] @ g[idx(y1)].conj().T
                Un2[i, mu] = g[i] @ Un[i, mu] @ g[idx(y3)].conj().T
    fat2 = _real(Uf2.reshape(V, 4, 9)).copy()
    lng2 = _real(Un2.reshape(V, 4, 9)).copy()
    psi = rng.standard_normal((Vh, 3)) + 1j * rng.standard_normal((Vh, 3))
    psi2 = np.einsum("sij,sj->si", g[Vh:], psi)  # source on odd sites
    out = _cplx(oracle.dslash(fat, lng, _real(psi).copy(), lat, 0))
    out2 = _cplx(oracle.dslash(fat2, lng2, _real(psi2).copy(), lat, 0))
    assert np.allclose(out2, np.einsum("sij,sj->si", g[:Vh], out), atol=1e-12)


@pytest.mark.parametrize("lat", [(4, 4, 4, 4), (2, 4, 2, 6)])
def test_dense_matrix_matches(lat):
    rng = np.random.default_rng(5)
    fat = _random_links(rng, lat, 0.9)
    lng = _random_links(rng, lat, 1 / 24)
    Vh = int(np.prod(lat)) // 2
    for p in (0, 1):
        M = dense_D(fat, lng, lat, p)
        src = rng.standard_normal((Vh, 3, 2))
        out = _cplx(oracle.dslash(fat, lng, src, lat, p)).reshape(-1)
        assert rel_err(_real(out), _real(M @ _cplx(src).reshape(-1))) < 1e-14
        # float path agrees to single precision
        out32 = oracle.dslash(fat.astype(np.float32), lng.astype(np.float32), src.astype(np.float32), lat, p)
        assert rel_err(out32.reshape(-1, 2), _real(out)) < 1e-6


def test_mdaggm_hermitian_psd():
    lat = (4, 4, 4, 4)
    rng = np.random.default_rng(6)
    fat = _random_links(rng, lat, 0.9)
    lng = _random_links(rng, lat, 1 / 24)
    K = -dense_D(fat, lng, lat, 0) @ dense_D(fat, lng, lat, 1)
    assert np.allclose(K, K.conj().T, atol=1e-13)
    assert np.linalg.eigvalsh(K).min() > -1e-12


@pytest.mark.parametrize("dtype,tolr", [(np.float64, 1e-9), (np.float32, 2e-4)])
def test_cgm_matches_dense_solves(dtype, tolr):
    lat = (4, 4, 4, 4)
    rng = np.random.default_rng(7)
    fat = _random_links(rng, lat, 0.9)
    lng = _random_links(rng, lat, 1 / 24)
    Vh = int(np.prod(lat)) // 2
    K = -dense_D(fat, lng, lat, 0) @ dense_D(fat, lng, lat, 1)
    b = rng.standard_normal((Vh, 3, 2))
    sig_abs = np.array([0.0025, 0.0125, 0.1025, 1.0025])
    shifts = np.concatenate([[sig_abs[0]], sig_abs[1:] - sig_abs[0]])
    tol = (1e-12 if dtype == np.float64 else 1e-6) ** 2 * float(np.sum(b * b))
    x, it = oracle.cgm(fat.astype(dtype), lng.astype(dtype), lat, 0, shifts.astype(dtype), b.astype(dtype), tol, 5000)
    assert 1 < it < 5000
    bc = _cplx(b).reshape(-1)
    for k, s in enumerate(sig_abs):
        xk = np.linalg.solve(K + s * np.eye(K.shape[0]), bc)
        assert rel_err(x[k], _real(xk).reshape(x[k].shape)) < tolr, k


@pytest.mark.parametrize("dtype,tolr", [(np.float64, 1e-9), (np.float32, 2e-4)])
def test_cgm_full_operator_matches_dense_solves(dtype, tolr):
    """parity -1: fields of both parities (even block first) and the full operator of a
    KUQCD_HISQ_DIRAC DiracHISQ, against dense solves of sigma_k - D D built from full-lattice
    coordinates (D D is block diagonal: -D_eo D_oe on the even block, -D_oe D_eo on the odd)."""
    lat = (4, 4, 4, 4)
    rng = np.random.default_rng(9)
    fat = _random_links(rng, lat, 0.9)
    lng = _random_links(rng, lat, 1 / 24)
    V = int(np.prod(lat))
    D0, D1 = dense_D(fat, lng, lat, 0), dense_D(fat, lng, lat, 1)
    n = D0.shape[0]
    K = np.zeros((2 * n, 2 * n), dtype=complex)
    K[:n, :n] = -D0 @ D1
    K[n:, n:] = -D1 @ D0
    b = rng.standard_normal((V, 3, 2))
    sig_abs = np.array([0.0025, 0.0125, 0.1025, 1.0025])
    shifts = np.concatenate([[sig_abs[0]], sig_abs[1:] - sig_abs[0]])
    tol = (1e-12 if dtype == np.float64 else 1e-6) ** 2 * float(np.sum(b * b))
    x, it = oracle.cgm(fat.astype(dtype), lng.astype(dtype), lat, -1, shifts.astype(dtype), b.astype(dtype), tol, 5000)
    assert 1 < it < 5000 and x[0].shape == b.shape
    bc = _cplx(b).reshape(-1)
    for k, s in enumerate(sig_abs):
        xk = np.linalg.solve(K + s * np.eye(2 * n), bc)
        assert rel_err(x[k], _real(xk).reshape(x[k].shape)) < tolr, k


def test_cgm_max_iter_and_drop_order():
    """max_iter bounds the iteration count; a converged heavy shift is dropped
    only from the end of the list (inverter_cgm.cpp:71-78)."""
    lat = (4, 4, 4, 4)
    rng = np.random.default_rng(8)
    fat = _random_links(rng, lat, 0.9)
    lng = _random_links(rng, lat, 1 / 24)
    Vh = int(np.prod(lat)) // 2
    b = rng.standard_normal((Vh, 3, 2))
    _, it = oracle.cgm(fat, lng, lat, 0, np.array([0.01, 0.5]), b, 1e-40, 7)
    assert it == 7
    _, it0 = oracle.cgm(fat, lng, lat, 0, np.array([0.01, 0.5]), b, 1e-3, 0)
    assert it0 == 1
