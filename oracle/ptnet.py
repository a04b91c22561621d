"""CPU restatement of the parallel-transport convolution preconditioner (TEST INFRASTRUCTURE ONLY).

numpy, FP64, on the reference host layouts. The network has no counterpart in
the reference; its taps are built from the same operator pieces as the
reference's hopping term (dirac_hisq.cpp:464-502: forward hop U(x) psi(x+h mu),
backward hop U(x-h mu)^+ psi(x-h mu), fat links with the 1/2 and long links with
c3000 = -1/48 of dirac_hisq.cpp:40), so that with W1[0][t] = d_t, W2[0][t] = -d_t
(d = +1 forward, -1 backward) and W0 = sigma it is exactly sigma - D D.

Pinning (tests/test_oracle_ptnet.py): that reduction against the C oracle's
mdaggm, gauge covariance, the adjoint relation between the taps of opposite
sign, the neighbour tables against the C oracle, and the gradient against
central finite differences of the loss.

Weights: complex, tap t = mu*4 + 2*(backward) + (hop 3); the kernels use
w_eff = bf16(float(w / nu_hop)) where nu_hop is the RMS row norm of the scaled
hop links.
"""
from __future__ import annotations

import numpy as np

from . import mdaggm as _mdaggm

NT = 16
NH = 4
C3000 = -1.0 / 48.0


def tap_index(mu: int, backward: bool, hop: int) -> int:
    return mu * 4 + (2 if backward else 0) + (1 if hop == 3 else 0)


def tap_of(t: int):
    """(mu, backward, hop) of tap t."""
    return t // 4, bool(t & 2), 3 if (t & 1) else 1


def flip(t: int) -> int:
    return t ^ 2


def tap_sign(t: int) -> float:
    """d_t of the hopping term D = sum_t d_t T_t."""
    return -1.0 if (t & 2) else 1.0


# ---------------------------------------------------------------- layouts
def links(fat: np.ndarray, lng: np.ndarray, lat):
    """Scaled complex links (2, Vh, 4, 3, 3) [parity, cb site, mu, row, col]: 1/2 fat, c3000 long."""
    V = int(np.prod(lat))
    f = np.asarray(fat, np.float64).reshape(2, V // 2, 4, 3, 3, 2)
    g = np.asarray(lng, np.float64).reshape(2, V // 2, 4, 3, 3, 2)
    return 0.5 * (f[..., 0] + 1j * f[..., 1]), C3000 * (g[..., 0] + 1j * g[..., 1])


def spinor(a: np.ndarray) -> np.ndarray:
    """reference host spinor (Vh*6 doubles) -> complex (Vh, 3)."""
    a = np.asarray(a, np.float64).reshape(-1, 3, 2)
    return a[..., 0] + 1j * a[..., 1]


def host(v: np.ndarray) -> np.ndarray:
    """complex (Vh, 3) -> reference host layout (Vh, 3, 2)."""
    return np.ascontiguousarray(np.stack([v.real, v.imag], axis=-1))


# ---------------------------------------------------------------- index math
def coords(lat, q: int):
    """full coordinates (x, y, z, t) of the checkerboard sites of parity q (index_utils.cpp:91-111)."""
    L0, L1, L2, _ = (int(v) for v in lat)
    Vh = int(np.prod(lat)) // 2
    lxh = L0 // 2
    s = np.arange(Vh)
    xcb = s % lxh
    r = s // lxh
    y = r % L1
    r //= L1
    z = r % L2
    t = r // L2
    x = 2 * xcb + ((y + z + t + q) & 1)
    return [x, y, z, t]


def nb(lat, q: int, mu: int, h: int) -> np.ndarray:
    """checkerboard index of the neighbour at hop h along mu (the wrapped full coordinate's x >> 1)."""
    L = [int(v) for v in lat]
    c = coords(lat, q)
    c[mu] = (c[mu] + h) % L[mu]
    return (c[0] >> 1) + (L[0] // 2) * (c[1] + L[1] * (c[2] + L[2] * c[3]))


class Geometry:
    """neighbour tables nb[qo][t] for every output parity and tap."""

    def __init__(self, lat):
        self.lat = tuple(int(v) for v in lat)
        self.Vh = int(np.prod(lat)) // 2
        self.nb = [[None] * NT for _ in range(2)]
        for qo in range(2):
            for t in range(NT):
                mu, back, hop = tap_of(t)
                self.nb[qo][t] = nb(lat, qo, mu, -hop if back else hop)


def tap(G: Geometry, U1, U3, qo: int, t: int, v: np.ndarray) -> np.ndarray:
    """(T_t v) on parity qo for v on parity 1-qo; v shape (..., Vh, 3)."""
    mu, back, hop = tap_of(t)
    U = U1 if hop == 1 else U3
    n = G.nb[qo][t]
    vn = v[..., n, :]
    if not back:
        return np.einsum("sij,...sj->...si", U[qo, :, mu], vn)
    return np.einsum("sji,...sj->...si", U[1 - qo, n, mu].conj(), vn)


# ---------------------------------------------------------------- weights
def bf16(x: np.ndarray) -> np.ndarray:
    """round float32 values to bfloat16 (nearest, ties to even), returned as float64."""
    shape = np.shape(x)
    f = np.array(x, np.float32).reshape(-1)
    u = f.view(np.uint32)
    finite = (u & 0x7F800000) != 0x7F800000
    r = (u + np.uint32(0x7FFF) + ((u >> 16) & 1)) & np.uint32(0xFFFF0000)
    u[finite] = r[finite]
    return f.astype(np.float64).reshape(shape)


def norms(U1, U3) -> tuple[float, float]:
    """RMS row norms of the scaled fat and long links."""
    n = U1.shape[1] * 8 * 3
    return float(np.sqrt(np.sum(np.abs(U1) ** 2) / n)), float(np.sqrt(np.sum(np.abs(U3) ** 2) / n))


def effective(W1, W2, w0, nu, round_bf16: bool = True):
    """kernel weights from master weights: bf16(float(w / nu_hop)); w0 unscaled."""
    s = np.array([1.0 / nu[1] if (t & 1) else 1.0 / nu[0] for t in range(NT)])

    def cvt(w):
        w = np.asarray(w, np.complex128)
        if not round_bf16:
            return w
        re = np.asarray(np.float32(w.real), np.float64)
        im = np.asarray(np.float32(w.imag), np.float64)
        return bf16(re) + 1j * bf16(im)

    def master(w):  # the kernels take the float32 master values
        w = np.asarray(w, np.complex128)
        if not round_bf16:
            return w
        return np.float32(w.real).astype(np.float64) + 1j * np.float32(w.imag).astype(np.float64)

    W1m, W2m, w0m = master(W1), master(W2), master(np.asarray(w0))
    return cvt(W1m * s), cvt(W2m * s), complex(cvt(w0m))


# ---------------------------------------------------------------- network
def forward(G: Geometry, U1, U3, W1e, W2e, w0e, q: int, r: np.ndarray):
    """h (4, Vh, 3) on parity 1-q and z (Vh, 3) on parity q, with effective weights."""
    T = [tap(G, U1, U3, 1 - q, t, r) for t in range(NT)]
    h = np.einsum("ct,tsi->csi", W1e, np.array(T))
    return h, layer2(G, U1, U3, W2e, w0e, q, r, h)


def layer2(G: Geometry, U1, U3, W2e, w0e, q: int, r: np.ndarray, h: np.ndarray):
    """z = w0 r + sum_t T_t(sum_c W2[c][t] h_c) on parity q from given hidden channels h (4, Vh, 3)."""
    z = w0e * r
    for t in range(NT):
        z = z + tap(G, U1, U3, q, t, np.einsum("c,csi->si", W2e[:, t], h))
    return z


def bf16_complex(h: np.ndarray) -> np.ndarray:
    """complex values with real and imaginary parts rounded to bfloat16 (from float32)."""
    return bf16(np.real(h)) + 1j * bf16(np.imag(h))


def backward(G: Geometry, U1, U3, W2e, q: int, r, h, g):
    """dL/dconj of the effective weights, given g = dL/dconj(z)."""
    gh = np.zeros_like(h)
    for t in range(NT):
        gh += np.conj(W2e[:, t])[:, None, None] * tap(G, U1, U3, 1 - q, flip(t), g)[None]
    gW1 = np.zeros((NH, NT), np.complex128)
    gW2 = np.zeros((NH, NT), np.complex128)
    for t in range(NT):
        Tr = tap(G, U1, U3, 1 - q, t, r)
        Tg = tap(G, U1, U3, 1 - q, flip(t), g)
        for c in range(NH):
            gW1[c, t] = np.vdot(Tr, gh[c])
            gW2[c, t] = np.vdot(h[c], Tg)
    gw0 = np.vdot(r, g)
    return gW1, gW2, gw0


def apply_A(fat, lng, lat, q: int, sigma: float, z: np.ndarray) -> np.ndarray:
    """(sigma - D D) z with the C oracle's hopping term."""
    dd = _mdaggm(np.ascontiguousarray(fat, np.float64), np.ascontiguousarray(lng, np.float64),
                 host(z).ravel(), lat, q)
    return sigma * z - spinor(dd)


def loss_grad(fat, lng, lat, q: int, sigma: float, W1, W2, w0, r, e, nu=None, round_bf16=True, want_grad=True):
    """L = |A(z - e)|^2 / |A e|^2 with z = M r (the error of the correction z in the A^2 norm; for a
    solver pair r = A e it is the residual left by one application of M), and dL/dconj of the master
    weights (complex [129])."""
    U1, U3 = links(fat, lng, lat)
    G = Geometry(lat)
    nu = norms(U1, U3) if nu is None else nu
    W1e, W2e, w0e = effective(W1, W2, w0, nu, round_bf16)
    r = spinor(r) if np.asarray(r).dtype != np.complex128 else r
    e = spinor(e) if np.asarray(e).dtype != np.complex128 else e
    h, z = forward(G, U1, U3, W1e, W2e, w0e, q, r)
    Az = apply_A(fat, lng, lat, q, sigma, z)
    Ae = apply_A(fat, lng, lat, q, sigma, e)
    d = Az - Ae
    n2 = np.vdot(Ae, Ae).real
    L = np.vdot(d, d).real / n2
    if not want_grad:
        return L, None
    g = apply_A(fat, lng, lat, q, sigma, d) / n2
    gW1, gW2, gw0 = backward(G, U1, U3, W2e, q, r, h, g)
    s = np.array([1.0 / nu[1] if (t & 1) else 1.0 / nu[0] for t in range(NT)])
    return L, np.concatenate([(gW1 * s).ravel(), (gW2 * s).ravel(), [gw0]])


def apply_layer2(fat, lng, lat, q: int, W2, w0, r, h, nu=None, round_bf16=True):
    """z of the network's second layer from given hidden channels h (4, Vh, 3), master weights."""
    U1, U3 = links(fat, lng, lat)
    G = Geometry(lat)
    nu = norms(U1, U3) if nu is None else nu
    _, W2e, w0e = effective(np.zeros_like(np.asarray(W2)), W2, w0, nu, round_bf16)
    return layer2(G, U1, U3, W2e, w0e, q, spinor(r), np.asarray(h, np.complex128))


def apply_net(fat, lng, lat, q: int, W1, W2, w0, r, nu=None, round_bf16=True):
    """(h, z) of the network on reference-layout input r, master weights."""
    U1, U3 = links(fat, lng, lat)
    G = Geometry(lat)
    nu = norms(U1, U3) if nu is None else nu
    W1e, W2e, w0e = effective(W1, W2, w0, nu, round_bf16)
    return forward(G, U1, U3, W1e, W2e, w0e, q, spinor(r))
