"""The mixed solver's deferred solution updates (segment folds), checked on the host: the
multishift recurrence of inverter_cgm.cpp:42-78 runs on a small dense SPD operator with the
reference's per-iteration updates x_k -= B_k p_k, p_k = Z_k r + a_k p_k, while the same
iteration is recorded (B_k, Z_k, a_k, shifts updated, residuals) as the GPU runner records it.
The library's coefficient routine (the function the fold kernel runs, through
kq_selftest_seg_coef) then rebuilds every x_k and p_k from the segment's start vectors and the
recorded residuals; the two agree to rounding, over segments of 1 to 32 iterations, with shifts
dropped mid-segment, and with the diverged-segment rule (keep_x = 0 drops the x updates)."""
import ctypes as C

import numpy as np
import pytest

KH, KS = 32, 32


def _coef(kq, h, J, k, keep_x=True):
    out = np.zeros(2 * (KH + 1))
    kq.L.lib().kq_selftest_seg_coef(h.ctypes.data, J, k, int(keep_x), out.ctypes.data)
    return out


def _problem(n=48, seed=0):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    K = q @ np.diag(np.logspace(-3, 1, n)) @ q.T
    b = rng.normal(size=n)
    sig = np.array([1e-3, 3e-3, 1e-2, 0.1, 1.0])
    return K, b, np.concatenate([[sig[0]], sig[1:] - sig[0]])


@pytest.mark.parametrize("seg_len,drops", [(1, False), (7, False), (16, True), (32, True), (9, True)])
def test_segment_fold_matches_per_iteration_updates(kq, seg_len, drops):
    K, b, shifts = _problem()
    ns, n = len(shifts), len(b)
    a, B, Z, Zm1 = np.zeros(ns), np.ones(ns), np.ones(ns), np.ones(ns)
    r = b.copy()
    p = [b.copy() for _ in range(ns)]
    x = [np.zeros(n) for _ in range(ns)]
    xd = [np.zeros(n) for _ in range(ns)]
    pd = [b.copy() for _ in range(ns)]
    p0_start = b.copy()
    h = np.zeros((KH, 4, KS))
    rec = []
    rr, Bm1, live = r @ r, 1.0, ns
    fold_live = ns  # shifts with updates since the segment start

    def fold():
        nonlocal rec, p0_start, fold_live
        J = len(rec)
        for k in range(fold_live):
            c = _coef(kq, h, J, k)
            start = p0_start if k == 0 else pd[k]
            xd[k] = xd[k] + c[0] * start + sum(c[1 + j] * rec[j] for j in range(J))
            if 1 <= k < live:
                pd[k] = c[KH + 1] * start + sum(c[KH + 2 + j] * rec[j] for j in range(J))
        p0_start = pd[0].copy()
        rec = []
        fold_live = live

    for it in range(70):
        s = shifts[0] * pd[0] + K @ pd[0]
        B[0] = -rr / (pd[0] @ s)
        for j in range(1, live):
            rr1 = Bm1 * Zm1[j] / (B[0] * a[0] * (Zm1[j] - Z[j]) + Zm1[j] * Bm1 * (1.0 - shifts[j] * B[0]))
            Zm1[j] = Z[j]
            Z[j] *= rr1
            B[j] = B[0] * rr1
        Bm1 = B[0]
        r = r + B[0] * s
        l2 = r @ r
        a[0] = l2 / rr
        rr = l2
        for j in range(1, live):
            a[j] = a[0] * Z[j] * B[j] / (Zm1[j] * B[0])
        jj = len(rec)
        h[jj, 0, :live], h[jj, 1, :live], h[jj, 2, :live] = B[:live], Z[:live], a[:live]
        h[jj, 3, 0] = live
        rec.append(r.copy())
        for k in range(live):
            x[k] = x[k] - B[k] * p[k]
            p[k] = Z[k] * r + a[k] * p[k]
        pd[0] = Z[0] * r + a[0] * pd[0]  # the runner keeps p_0 current every iteration
        if drops and it % 11 == 10 and live > 1:
            live -= 1
        if len(rec) == seg_len:
            fold()
    fold()
    for k in range(ns):
        assert np.linalg.norm(x[k] - xd[k]) <= 1e-12 * np.linalg.norm(x[k]), k
        if k < live:
            assert np.linalg.norm(p[k] - pd[k]) <= 1e-12 * np.linalg.norm(p[k]), k
    # the deferred x_k solve the shifted systems as well as the per-iteration ones
    for k in range(ns):
        sig = shifts[0] + (shifts[k] if k else 0.0)
        res = np.linalg.norm(b - (sig * xd[k] + K @ xd[k])) / np.linalg.norm(b)
        ref = np.linalg.norm(b - (sig * x[k] + K @ x[k])) / np.linalg.norm(b)
        assert res <= 2 * ref + 1e-13, (k, res, ref)


def test_diverged_segment_keeps_no_x_update(kq):
    h = np.zeros((KH, 4, KS))
    h[:5, 0, 0], h[:5, 1, 0], h[:5, 2, 0], h[:5, 3, 0] = -0.5, 1.0, 0.3, 1
    kept = _coef(kq, h, 5, 0, keep_x=True)
    dropped = _coef(kq, h, 5, 0, keep_x=False)
    assert np.any(kept[:KH + 1] != 0) and not np.any(dropped[:KH + 1])
    # the search direction is rebuilt either way
    assert np.array_equal(kept[KH + 1:], dropped[KH + 1:])
    # a shift not updated in the segment gets nothing
    none = _coef(kq, h, 5, 3)
    assert not np.any(none[:KH + 1]) and none[KH + 1] == 1.0 and not np.any(none[KH + 2:])
