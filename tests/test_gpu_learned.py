"""Learned preconditioner on the GPU: the HIP forward, backward and weight-gradient
kernels against the CPU restatement (oracle/ptnet.py), the flexible CG against
the oracle's operator, the plain-CG fallback, training on solver-generated
pairs, and what the trained network gains over plain CG on held-out gauge fields.
"""
import numpy as np
import pytest

import oracle
import oracle.ptnet as P
from conftest import rel_err

pytestmark = pytest.mark.gpu


def _load_weak(kq, lat, seed, eps=0.3):
    fat, lng = kq.gen_gauge(lat, kind=kq.KQ_GEN_WEAK, eps=eps, fat_scale=0.9, seed_fat=2 * seed + 1,
                            seed_long=2 * seed + 2)
    kq.load_gauge_pair(fat, lng, lat)
    return fat.reshape(-1), lng.reshape(-1)


def _weights(kq, seed, scale=0.3):
    rng = np.random.default_rng(seed)
    W1 = scale * (rng.standard_normal((4, 16)) + 1j * rng.standard_normal((4, 16)))
    W2 = scale * (rng.standard_normal((4, 16)) + 1j * rng.standard_normal((4, 16)))
    w0 = 0.8 + 0.3j
    return kq.weights_from_array(W1, W2, w0)


def _c(a):
    a = np.asarray(a).reshape(-1, 2)
    return a[:, 0] + 1j * a[:, 1]


def _rel(a, b):
    a, b = np.asarray(a).reshape(-1), np.asarray(b).reshape(-1)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("lat,q", [((8, 8, 8, 8), 0), ((6, 4, 4, 6), 1), ((4, 6, 2, 6), 0)])
def test_forward_matches_oracle(gpu, lat, q):
    fat, lng = _load_weak(gpu, lat, 1)
    nu = gpu.ptnet_norms()
    nu_o = P.norms(*P.links(fat, lng, lat))
    assert abs(nu[0] - nu_o[0]) < 1e-12 * nu_o[0] and abs(nu[1] - nu_o[1]) < 1e-12 * nu_o[1]
    w = _weights(gpu, 5)
    W1, W2, w0 = gpu.weights_to_array(w)
    r = gpu.gen_source(lat, parity=q, seed=9)
    h_o, z_o = P.apply_net(fat, lng, lat, q, W1, W2, w0, r, nu=nu)
    z, h = gpu.ptnet_forward(w, r, parity=q, link_format=gpu.KQ_MODE_MIXED_FP32, hidden=True)
    # FP32 links and accumulation against FP64 (same bfloat16 weights)
    assert _rel(_c(z), z_o) < 1e-5
    for c in range(4):
        assert _rel(_c(h[c]), h_o[c]) < 1e-5
    z16 = gpu.ptnet_forward(w, r, parity=q, link_format=gpu.KQ_MODE_MIXED_FP16)
    assert _rel(_c(z16), z_o) < 5e-3


@pytest.mark.parametrize("lat,q", [((8, 8, 8, 8), 0), ((6, 4, 4, 6), 1)])
def test_forward_bf16_hidden_matches_oracle(gpu, lat, q):
    """The flexible CG's form of the network: hidden channels rounded to bfloat16 between the layers."""
    fat, lng = _load_weak(gpu, lat, 1)
    nu = gpu.ptnet_norms()
    w = _weights(gpu, 5)
    W1, W2, w0 = gpu.weights_to_array(w)
    r = gpu.gen_source(lat, parity=q, seed=9)
    h_o, z_o = P.apply_net(fat, lng, lat, q, W1, W2, w0, r, nu=nu)
    z16, h16 = gpu.ptnet_forward(w, r, parity=q, link_format=gpu.KQ_MODE_MIXED_FP32, hidden=True, hidden16=True)
    hc = np.array([_c(h16[c]).reshape(-1, 3) for c in range(4)])
    # every hidden entry is a bfloat16 number within half a bfloat16 step (2^-8 relative) of the
    # FP64 first layer (plus the FP32 accumulation error)
    assert np.array_equal(P.bf16_complex(hc), hc)
    for c in range(4):
        d = hc[c] - h_o[c]
        tiny = 1e-6 * np.abs(h_o[c]).max()
        assert np.all(np.abs(d.real) <= 2.0 ** -8 * np.abs(h_o[c].real) + tiny)
        assert np.all(np.abs(d.imag) <= 2.0 ** -8 * np.abs(h_o[c].imag) + tiny)
    # the second layer on exactly these channels is the oracle's, to FP32 accuracy
    z_o2 = P.apply_layer2(fat, lng, lat, q, W2, w0, r, hc, nu=nu)
    assert _rel(_c(z16), z_o2) < 1e-5
    # the rounding moves z by about the bfloat16 step, no more
    assert 1e-5 < _rel(_c(z16), z_o) < 1e-2


@pytest.mark.parametrize("lat,q", [((8, 8, 8, 8), 0), ((6, 4, 4, 6), 1)])
def test_loss_and_gradient_match_oracle(gpu, lat, q):
    fat, lng = _load_weak(gpu, lat, 2)
    nu = gpu.ptnet_norms()
    w = _weights(gpu, 6)
    W1, W2, w0 = gpu.weights_to_array(w)
    rng = np.random.default_rng(3)
    r = gpu.gen_source(lat, parity=q, seed=10)
    e = r + 0.3 * rng.standard_normal(r.shape)
    sigma = 0.02
    L_o, g_o = P.loss_grad(fat, lng, lat, q, sigma, W1, W2, w0, r, e, nu=nu)
    L, g = gpu.ptnet_loss_grad(w, r, e, sigma, parity=q)
    assert L_o > 0.0
    assert abs(L - L_o) < 1e-5 * L_o
    assert np.linalg.norm(g - g_o) < 1e-4 * np.linalg.norm(g_o)
    # per block: hop-1 taps, hop-3 taps (scaled by 1/nu_long), skip weight
    for sel in (slice(0, 128, 2), slice(1, 128, 2), slice(128, 129)):
        assert np.linalg.norm(g[sel] - g_o[sel]) < 1e-4 * np.linalg.norm(g_o[sel])


def _true_rel(fat, lng, lat, q, sigma, b, x):
    Ax = P.apply_A(fat, lng, lat, q, sigma, _c(x).reshape(-1, 3))
    return np.linalg.norm(_c(b).reshape(-1, 3) - Ax) / np.linalg.norm(_c(b))


@pytest.mark.parametrize("mode", [2, 3])
def test_fcg_plain_reaches_tolerance(gpu, mode):
    lat, q = (8, 8, 8, 8), 0
    fat, lng = _load_weak(gpu, lat, 3)
    sigma = 0.01 ** 2 + 1e-4
    b = gpu.gen_source(lat, parity=q, seed=11)
    x = np.zeros_like(b)
    st, rep = gpu.fcg_refine(b, x, sigma, parity=q, opt=gpu.refine_options(learned=0, rel_tol=1e-10, mode=mode))
    assert st == gpu.KUQCD_SUCCESS
    assert rep.converged == 1 and rep.fallback == 0 and rep.iter_learned == 0
    assert rep.true_rel_res <= 1e-10
    assert _true_rel(fat, lng, lat, q, sigma, b, x) <= 1.05e-10
    xs, _ = oracle.cgm(fat, lng, lat, q, np.array([sigma]), b.reshape(-1), 1e-26 * np.sum(b * b), 100000)
    assert rel_err(x.reshape(-1), xs[0]) < 1e-7


def test_fcg_falls_back_when_network_fails(gpu):
    """A network that returns zero makes no progress: the refinement switches to plain CG."""
    lat, q = (8, 8, 8, 8), 1
    fat, lng = _load_weak(gpu, lat, 4)
    sigma = 0.01 ** 2 + 1e-4
    gpu.ptnet_set_weights(gpu.weights_from_array(np.zeros((4, 16)), np.zeros((4, 16)), 0.0))
    b = gpu.gen_source(lat, parity=q, seed=12)
    x = np.zeros_like(b)
    st, rep = gpu.fcg_refine(b, x, sigma, parity=q,
                             opt=gpu.refine_options(learned=1, rel_tol=1e-10, stall_iter=40))
    assert st == gpu.KUQCD_SUCCESS
    assert rep.fallback == 1 and rep.converged == 1
    assert rep.iter_learned == 40
    assert _true_rel(fat, lng, lat, q, sigma, b, x) <= 1.05e-10


def test_identity_network_is_plain_cg(gpu):
    """W1 = W2 = 0, W0 = 1 makes M the identity: same iterations as plain CG (up to rounding)."""
    lat, q = (8, 8, 8, 8), 0
    _load_weak(gpu, lat, 5)
    sigma = 0.01 ** 2 + 1e-4
    gpu.ptnet_set_weights(gpu.weights_from_array(np.zeros((4, 16)), np.zeros((4, 16)), 1.0))
    b = gpu.gen_source(lat, parity=q, seed=13)
    xa, xb = np.zeros_like(b), np.zeros_like(b)
    _, ra = gpu.fcg_refine(b, xa, sigma, parity=q, opt=gpu.refine_options(learned=1, rel_tol=1e-10))
    _, rb = gpu.fcg_refine(b, xb, sigma, parity=q, opt=gpu.refine_options(learned=0, rel_tol=1e-10))
    assert ra.converged and rb.converged and ra.fallback == 0
    assert abs(ra.iter - rb.iter) <= max(2, rb.iter // 50)


def test_training_reduces_loss_and_refines_held_out(gpu):
    """Solver-generated pairs on 3 gauge fields, Adam on the GPU kernels, then the flexible CG
    with the trained network on a held-out field: the training loss falls, the refinement from
    1e-6 reaches 1e-10 (FP64 true residual) without falling back, and the iteration counts of
    the learned and the plain refinement are reported."""
    lat, q = (16, 16, 16, 16), 0
    mass = 0.005
    sigma = mass * mass + 1e-4
    gpu.learned_clear()
    for s in range(3):
        _load_weak(gpu, lat, 100 + s)
        n = gpu.learned_collect(q, sigma, nsrc=1, seed=1000 + s, rel_tol=1e-11, min_level=1e-8)
    assert n >= 3 * 6
    rep = gpu.learned_train(steps=150, batch=8, seed=1)
    assert rep.nsamples == n and rep.ngauges == 3
    assert rep.loss_final < rep.loss_initial
    gpu.learned_clear()
    fat, lng = _load_weak(gpu, lat, 500)
    b = gpu.gen_source(lat, parity=q, seed=7000)
    x = np.zeros_like(b)
    st, r6 = gpu.fcg_refine(b, x, sigma, parity=q, opt=gpu.refine_options(learned=0, rel_tol=1e-6))
    assert st == gpu.KUQCD_SUCCESS and r6.true_rel_res <= 1e-6
    res = {}
    for name, lf in (("plain", 0), ("learned", 1)):
        xx = x.copy()
        st, r = gpu.fcg_refine(b, xx, sigma, parity=q, opt=gpu.refine_options(learned=lf, rel_tol=1e-10))
        assert st == gpu.KUQCD_SUCCESS and r.converged == 1
        assert _true_rel(fat, lng, lat, q, sigma, b, xx) <= 1.05e-10
        res[name] = r
    assert res["learned"].fallback == 0 and res["learned"].iter_learned == res["learned"].iter
    print(f"\nloss {rep.loss_initial:.4f} -> {rep.loss_final:.4f}; refinement 1e-6 -> 1e-10: plain "
          f"{res['plain'].iter} iterations {res['plain'].seconds:.3f}s, learned {res['learned'].iter} iterations "
          f"{res['learned'].seconds:.3f}s")


def test_mixed_solve_with_learned_refinement(gpu):
    lat, q = (16, 16, 16, 16), 0
    mass = 0.01
    shifts = np.concatenate([[mass * mass + 1e-4], np.logspace(-4, 0, 6)[1:] - 1e-4])
    _load_weak(gpu, lat, 7)
    gpu.learned_clear()
    gpu.learned_collect(q, shifts[0], nsrc=1, seed=5, rel_tol=1e-11, min_level=1e-8)
    gpu.learned_train(steps=60, batch=4, seed=2)
    gpu.learned_clear()
    b = gpu.gen_source(lat, parity=q, seed=77)
    for mode in (2, 3):
        x = [np.zeros_like(b) for _ in shifts]
        p = gpu.InvertParam(shifts=shifts, tol=1e-20, max_iter=20000)
        st, rep = gpu.solve_ex(x, b, p, gpu.default_options(mode=mode, rel_tol=1e-10, learned=1, verbose=0))
        assert st == gpu.KUQCD_SUCCESS, gpu.last_error()
        assert rep.learned_used + rep.learned_fallback == 1
        assert rep.refine_iter[0] > 0
        assert max(rep.true_rel_res[:len(shifts)]) <= 1e-10
