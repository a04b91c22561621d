"""Pins the CPU restatement of the parallel-transport preconditioner (oracle/ptnet.py).

  * neighbour tables equal the C oracle's index math (odd and even extents)
  * with W1[0][t] = d_t, W2[0][t] = -d_t, W0 = sigma the network is sigma - D D,
    checked against the C oracle's mdaggm (a wrong link site, missing dagger,
    wrong hop length or sign in any tap fails this)
  * gauge covariance with random weights (no tap may mix links and sites wrongly)
  * the adjoint of tap t is the tap of opposite sign (used by the backward pass)
  * the gradient equals central finite differences of the loss, along random
    complex directions and along single real/imaginary coordinates
  * bfloat16 rounding: nearest, ties to even
"""
import numpy as np
import pytest

import oracle
import oracle.ptnet as P
from conftest import rel_err
from test_oracle import _cb, _cplx, _haar, _random_links, _real, _sites


def _rand_spinor(rng, Vh):
    return rng.standard_normal((Vh, 3)) + 1j * rng.standard_normal((Vh, 3))


@pytest.mark.parametrize("lat", [(4, 6, 4, 4), (6, 3, 2, 5)])
def test_nb_tables_match_c_oracle(lat):
    Vh = int(np.prod(lat)) // 2
    for q in (0, 1):
        for mu in range(4):
            for h in (1, 3, -1, -3):
                t = P.nb(lat, q, mu, h)
                ref = np.array([oracle.nb(lat, q, s, mu, h) for s in range(Vh)])
                assert np.array_equal(t, ref), (q, mu, h)


@pytest.mark.parametrize("lat", [(4, 4, 4, 4), (4, 6, 2, 6), (6, 3, 2, 5)])
@pytest.mark.parametrize("q", [0, 1])
def test_network_reduces_to_operator(lat, q):
    rng = np.random.default_rng(11)
    fat = _random_links(rng, lat, 0.9)
    lng = _random_links(rng, lat, 1 / 24)
    Vh = int(np.prod(lat)) // 2
    sigma = 0.37
    U1, U3 = P.links(fat, lng, lat)
    G = P.Geometry(lat)
    d = np.array([P.tap_sign(t) for t in range(P.NT)])
    W1 = np.zeros((P.NH, P.NT), complex)
    W2 = np.zeros((P.NH, P.NT), complex)
    W1[0] = d
    W2[0] = -d
    r = _rand_spinor(rng, Vh)
    h, z = P.forward(G, U1, U3, W1, W2, sigma, q, r)
    ref = sigma * r - _cplx(oracle.mdaggm(fat, lng, _real(r).copy(), lat, q))
    assert rel_err(np.stack([z.real, z.imag]), np.stack([ref.real, ref.imag])) < 1e-13
    # hidden channel 0 is D r
    Dr = _cplx(oracle.dslash(fat, lng, _real(r).copy(), lat, 1 - q))
    assert rel_err(np.stack([h[0].real, h[0].imag]), np.stack([Dr.real, Dr.imag])) < 1e-13
    assert np.all(h[1:] == 0)


def test_gauge_covariance():
    lat = (4, 6, 4, 4)
    V = int(np.prod(lat))
    Vh = V // 2
    rng = np.random.default_rng(12)
    fat = _random_links(rng, lat, 0.9)
    lng = _random_links(rng, lat, 1 / 24)
    g = _haar(rng, V)

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
    W1 = rng.standard_normal((P.NH, P.NT)) + 1j * rng.standard_normal((P.NH, P.NT))
    W2 = rng.standard_normal((P.NH, P.NT)) + 1j * rng.standard_normal((P.NH, P.NT))
    w0 = 0.7 - 0.2j
    q = 0
    r = _rand_spinor(rng, Vh)
    G = P.Geometry(lat)
    h, z = P.forward(G, *P.links(fat, lng, lat), W1, W2, w0, q, r)
    r2 = np.einsum("sij,sj->si", g[:Vh], r)
    h2, z2 = P.forward(G, *P.links(fat2, lng2, lat), W1, W2, w0, q, r2)
    assert np.allclose(z2, np.einsum("sij,sj->si", g[:Vh], z), atol=1e-11)
    assert np.allclose(h2, np.einsum("sij,csj->csi", g[Vh:], h), atol=1e-11)


def test_tap_adjoint():
    lat = (4, 4, 6, 4)
    Vh = int(np.prod(lat)) // 2
    rng = np.random.default_rng(13)
    fat = _random_links(rng, lat, 0.9)
    lng = _random_links(rng, lat, 1 / 24)
    U1, U3 = P.links(fat, lng, lat)
    G = P.Geometry(lat)
    for qo in (0, 1):
        a = _rand_spinor(rng, Vh)  # parity 1-qo
        b = _rand_spinor(rng, Vh)  # parity qo
        for t in range(P.NT):
            lhs = np.vdot(P.tap(G, U1, U3, qo, t, a), b)
            rhs = np.vdot(a, P.tap(G, U1, U3, 1 - qo, P.flip(t), b))
            assert abs(lhs - rhs) < 1e-12 * abs(lhs), (qo, t)


def test_gradient_finite_differences():
    lat = (4, 4, 4, 6)
    Vh = int(np.prod(lat)) // 2
    rng = np.random.default_rng(14)
    fat = _random_links(rng, lat, 0.9)
    lng = _random_links(rng, lat, 1 / 24)
    q, sigma = 1, 0.05
    U1, U3 = P.links(fat, lng, lat)
    nu = P.norms(U1, U3)
    W1 = 0.3 * (rng.standard_normal((P.NH, P.NT)) + 1j * rng.standard_normal((P.NH, P.NT)))
    W2 = 0.3 * (rng.standard_normal((P.NH, P.NT)) + 1j * rng.standard_normal((P.NH, P.NT)))
    w0 = 1.0 + 0.1j
    r = _rand_spinor(rng, Vh)
    e = r + 0.2 * _rand_spinor(rng, Vh)

    def L(v):
        return P.loss_grad(fat, lng, lat, q, sigma, v[:64].reshape(4, 16), v[64:128].reshape(4, 16), v[128], r, e,
                           nu=nu, round_bf16=False, want_grad=False)[0]

    w = np.concatenate([W1.ravel(), W2.ravel(), [w0]])
    L0, grad = P.loss_grad(fat, lng, lat, q, sigma, W1, W2, w0, r, e, nu=nu, round_bf16=False)
    assert L0 > 0.0
    assert abs(L(w) - L0) < 1e-14
    eps = 1e-6
    for k in range(4):
        dw = rng.standard_normal(w.size) + 1j * rng.standard_normal(w.size)
        fd = (L(w + eps * dw) - L(w - eps * dw)) / (2 * eps)
        pred = 2.0 * np.vdot(grad, dw).real
        assert abs(fd - pred) < 1e-6 * max(1.0, abs(pred)), (k, fd, pred)
    # single coordinates: real and imaginary parts of hop-1, hop-3 and skip weights
    for j in (0, 1, 5, 66, 71, 128):
        for unit in (1.0, 1j):
            dw = np.zeros(w.size, complex)
            dw[j] = unit
            fd = (L(w + eps * dw) - L(w - eps * dw)) / (2 * eps)
            pred = 2.0 * (grad[j].real if unit == 1.0 else grad[j].imag)
            assert abs(fd - pred) < 1e-7 + 1e-5 * abs(pred), (j, unit, fd, pred)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -8 + 2.0 ** -20, -3.14159,
                  np.inf, -np.inf], np.float64)
    y = P.bf16(x)
    assert list(y[:5]) == [1.0, 1.0, 1.0 + 4 * 2.0 ** -8, 1.0 + 2 * 2.0 ** -8, -3.140625]
    assert np.isinf(y[5]) and np.isinf(y[6])
    assert np.isnan(P.bf16(np.array([np.nan]))[0])
