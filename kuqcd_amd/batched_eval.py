"""Time the multi-right-hand-side solve against single solves.

    python -m kuqcd_amd.batched_eval [--lattice 32] [--n 4,8,16] [--mode 2] [--reps 2]

Workload: one seeded weak-field gauge pair (the 32^4 generator of bench.py), N Gaussian
even-site sources (seeds 1000 + r), each with its own 12 shifts drawn log-uniform in [1e-4, 1]
(m = 0.01), FP64 true relative residual 1e-10 per shift. For each N, one batched mixed solve
(kq_solve_multishift_batched) is timed against the N single solves of the same sources
(kq_solve_multishift_ex), both with host arrays, after one warm-up of each; every shift's
reported FP64 residual is checked. Prints one JSON object.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import api as A


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--lattice", type=int, default=32)
    ap.add_argument("--n", default="4,8,16")
    ap.add_argument("--mode", type=int, default=A.KQ_MODE_MIXED_FP32)
    ap.add_argument("--reps", type=int, default=2)
    a = ap.parse_args(argv)
    lat = (a.lattice,) * 4
    A.set_verbosity(0)
    assert A.kuqcd_init() == A.KUQCD_SUCCESS
    fat, lng = A.gen_gauge(lat, kind=A.KQ_GEN_WEAK, eps=0.3, fat_scale=0.9, seed_fat=1, seed_long=2)
    A.load_gauge_pair(fat, lng, lat)
    mass, ns = 0.01, 12
    out = {"lattice": list(lat), "mode": a.mode, "rows": []}
    for n in [int(v) for v in a.n.split(",")]:
        rng = np.random.default_rng(4 + n)
        bs = [A.gen_source(lat, 0, seed=1000 + r) for r in range(n)]
        sh = np.zeros((n, ns))
        for r in range(n):
            s = np.sort(10 ** rng.uniform(-4, 0, ns))
            sh[r] = np.concatenate([[mass * mass + s[0]], s[1:] - s[0]])
        opt = A.default_options(mode=a.mode, refine=1, rel_tol=1e-10, verbose=0)
        xs = [np.zeros_like(bs[0]) for _ in range(n * ns)]

        def batched():
            st, reps = A.solve_batched(xs, bs, A.InvertParam(shifts=sh, tol=0.0, max_iter=40000), opt)
            assert st == A.KUQCD_SUCCESS, A.last_error()
            assert max(max(rp.true_rel_res[:ns]) for rp in reps) <= 1e-10
            return reps

        def singles():
            for r in range(n):
                x = [np.zeros_like(bs[0]) for _ in range(ns)]
                st, rep = A.solve_ex(x, bs[r], A.InvertParam(shifts=sh[r], tol=0.0, max_iter=40000), opt)
                assert st == A.KUQCD_SUCCESS, A.last_error()
                assert max(rep.true_rel_res[:ns]) <= 1e-10

        batched()
        singles()
        assert A.kq_synchronize() == 0
        t = time.perf_counter()
        for _ in range(a.reps):
            batched()
        tb = (time.perf_counter() - t) / a.reps
        t = time.perf_counter()
        for _ in range(a.reps):
            singles()
        ts = (time.perf_counter() - t) / a.reps
        row = {"n": n, "batched_s": round(tb, 4), "singles_s": round(ts, 4), "batched_per_source_ms": round(1e3 * tb / n, 1),
               "single_per_source_ms": round(1e3 * ts / n, 1), "speedup": round(ts / tb, 3)}
        out["rows"].append(row)
        print(row, file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)
    A.kuqcd_gauge_free()
    return 0


if __name__ == "__main__":
    sys.exit(main())
