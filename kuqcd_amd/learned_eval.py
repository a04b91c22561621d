"""Train the parallel-transport preconditioner on seeded gauge fields and measure
what it gains over plain CG on held-out ones.

    python -m kuqcd_amd.learned_eval [--lattice 24] [--mass 0.005] [--train 16] [--test 8]
                                     [--steps 400] [--save weights.json]

Protocol: for each training seed, a weak-field gauge pair is generated and loaded,
and (residual, error) pairs are collected from an FP64-reliable CG on the lightest
shift sigma = m^2 + 1e-4 (kq_learned_collect). Adam then trains the network
(kq_learned_train). On each held-out seed a Gaussian source is solved by plain CG
to relative residual 1e-6 and that iterate is refined to 1e-10 twice, by the
flexible CG with the trained network and by plain CG (the same code with z = r);
both FP64 true residuals are checked. Prints one JSON object.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import api as A


def _load(lat, seed):
    f, g = A.gen_gauge(lat, kind=A.KQ_GEN_WEAK, eps=0.3, fat_scale=0.9, seed_fat=2 * seed + 1, seed_long=2 * seed + 2)
    A.load_gauge_pair(f, g, lat)


def save_weights(w, path):
    W1, W2, w0 = A.weights_to_array(w)
    c = np.concatenate([W1.ravel(), W2.ravel(), [w0]])
    with open(path, "w") as f:
        json.dump({"re": c.real.tolist(), "im": c.imag.tolist()}, f)


def load_weights(path):
    with open(path) as f:
        d = json.load(f)
    c = np.array(d["re"]) + 1j * np.array(d["im"])
    return A.weights_from_array(c[:64].reshape(4, 16), c[64:128].reshape(4, 16), c[128])


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--lattice", type=int, default=24)
    ap.add_argument("--mass", type=float, default=0.005)
    ap.add_argument("--train", type=int, default=16)
    ap.add_argument("--test", type=int, default=8)
    ap.add_argument("--nsrc", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=3e-4)
    ap.add_argument("--mode", type=int, default=A.KQ_MODE_MIXED_FP32, help="link format of A in the refinement")
    ap.add_argument("--save", default="")
    a = ap.parse_args(argv)
    lat = (a.lattice,) * 4
    sigma = a.mass ** 2 + 1e-4
    A.set_verbosity(0)
    assert A.kuqcd_init() == A.KUQCD_SUCCESS
    A.learned_clear()
    t0 = time.time()
    for s in range(a.train):
        _load(lat, 100 + s)
        n = A.learned_collect(0, sigma, nsrc=a.nsrc, seed=1000 + s, rel_tol=1e-11, min_level=1e-8)
        print(f"train seed {s}: {n} pairs so far", file=sys.stderr, flush=True)
    t_collect = time.time() - t0
    rep = A.learned_train(steps=a.steps, batch=a.batch, lr=a.lr, lr_final=a.lr / 10, seed=1, init=2)
    w = A.ptnet_get_weights()
    if a.save:
        save_weights(w, a.save)
    print(f"trained: loss {rep.loss_initial:.4f} -> {rep.loss_final:.4f} in {rep.seconds:.1f}s", file=sys.stderr,
          flush=True)
    rows = []
    for s in range(a.test):
        _load(lat, 500 + s)
        b = A.gen_source(lat, parity=0, seed=7000 + s)
        x = np.zeros_like(b)
        A.fcg_refine(b, x, sigma, opt=A.refine_options(learned=0, rel_tol=1e-6, mode=a.mode))
        row = {"seed": 500 + s}
        for name, lf in (("plain", 0), ("learned", 1)):
            xx = x.copy()
            st, r = A.fcg_refine(b, xx, sigma, opt=A.refine_options(learned=lf, rel_tol=1e-10, mode=a.mode))
            row[name] = {"status": st, "iter": r.iter, "iter_net": r.iter_learned, "fallback": r.fallback,
                         "seconds": round(r.seconds, 4), "rel0": r.rel_res_initial, "rel": r.true_rel_res}
        rows.append(row)
        print(f"test seed {s}: plain {row['plain']['iter']} it {row['plain']['seconds']:.3f}s, learned "
              f"{row['learned']['iter']} it {row['learned']['seconds']:.3f}s fb {row['learned']['fallback']}",
              file=sys.stderr, flush=True)
    it_p = sum(r["plain"]["iter"] for r in rows)
    it_l = sum(r["learned"]["iter"] for r in rows)
    t_p = sum(r["plain"]["seconds"] for r in rows)
    t_l = sum(r["learned"]["seconds"] for r in rows)
    out = {
        "lattice": list(lat), "mass": a.mass, "sigma": sigma, "train_seeds": a.train, "test_seeds": a.test,
        "pairs": rep.nsamples, "collect_seconds": round(t_collect, 1), "train_seconds": round(rep.seconds, 1),
        "loss_initial": rep.loss_initial, "loss_final": rep.loss_final,
        "iteration_gain": round(it_p / max(1, it_l), 3), "time_gain": round(t_p / max(1e-9, t_l), 3),
        "all_converged": all(r[k]["rel"] <= 1e-10 for r in rows for k in ("plain", "learned")),
        "fallbacks": sum(r["learned"]["fallback"] for r in rows), "per_seed": rows,
    }
    print(json.dumps(out))
    A.kuqcd_gauge_free()
    A.kuqcd_end()
    return 0


if __name__ == "__main__":
    sys.exit(main())
