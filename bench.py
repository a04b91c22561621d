#!/usr/bin/env python3
"""Benchmark of the HISQ even-odd multishift CG on one MI355X.

Workload (synthetic, seeded): 32^4 periodic lattice, fat links 0.9 * exp(i 0.3 H)
and long links -1/24 * W(x)W(x+mu)W(x+2mu) with W = exp(i 0.3 H), staggered
phases folded in, Gaussian even-site source, mass m = 0.01, 12 shifts
log-spaced in [1e-4, 1] (absolute shifts m^2 + s_k), FP64 true relative
residual <= 1e-10 for every shift.

One step = one complete mixed-precision multishift solve as the 32^4 workload
specifies it: FP32 links and vectors with FP32 accumulation in the iteration,
FP64 reliable updates whenever |r| has dropped to 10 % of its value at the last
one (delta 0.1), FP64 recurrence scalars, per-shift FP64 verification and
refinement, with the source and the solutions resident in HBM. The line also
carries the plain FP64 GPU multishift CG time on the same problem, the same
headline solve with the source and solutions in host memory (PCIe transfers
included), the tuned variants (library default delta 0.02; FP32 fat with 16-bit
fixed-point long links; the headline mode with the learned-preconditioner
refinement of shift 0), the FP64 hopping-term roofline, and the oracle CGM run
to completion on the host cores (OMP_PROC_BIND=close, OMP_PLACES=cores), whose
solution also gives each shift's error of the mixed and FP64 solves.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "12-shift CGM time to rel. residual 1e-10, 32^4: mixed vs FP64; dslash %HBM peak"
HBM_PEAK_GBS = 8000.0
# arithmetic of each mixed mode: link storage + vectors + reliable updates
MODE_DTYPE = {2: "fp32+fp64", 3: "int16-fixed+fp32+fp64", 4: "fp32+int16-fixed(long links)+fp64"}
MODE_NAME = {2: "mixed: FP32 links", 3: "mixed: 16-bit links", 4: "mixed: FP32 fat + 16-bit long links"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pmc_traffic(path: Path, kernel: str):
    """HBM bytes per launch of a hopping-term kernel (name prefix as rocprofv3 prints it) from the committed
    rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor)."""
    if not path.exists():
        return None
    for block in path.read_text().split("\nkernel: ")[1:]:
        if block.startswith(kernel):
            m = re.search(r"traffic_bytes_per_launch\s*[:=]\s*([0-9.eE+]+)", block)
            return float(m.group(1)) if m else None
    return None


def host_cpu() -> dict:
    """CPU model, physical cores (lscpu) and the CPUs this process may run on."""
    info = {"model": "unknown CPU", "physical_cores": None, "cpus_allowed": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = dict(ln.split(":", 1) for ln in out.splitlines() if ":" in ln)
        kv = {k.strip(): v.strip() for k, v in kv.items()}
        info["model"] = kv.get("Model name", info["model"])
        if "Core(s) per socket" in kv and "Socket(s)" in kv:
            info["physical_cores"] = int(kv["Core(s) per socket"]) * int(kv["Socket(s)"])
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return info


def main() -> int:
    # the oracle's OpenMP threads: one per core, packed (read when its library is loaded)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    hc = host_cpu()  # before any OpenMP runtime binds this thread to one place
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--lattice", default="32,32,32,32")
    ap.add_argument("--nshift", type=int, default=12)
    ap.add_argument("--mass", type=float, default=0.01)
    ap.add_argument("--shift-min", type=float, default=1e-4, help="smallest of the log-spaced shifts (largest 1)")
    ap.add_argument("--rel-tol", type=float, default=1e-10)
    ap.add_argument("--mode", type=int, default=2,
                    help="mixed mode: 2 = FP32 links, 3 = 16-bit links, 4 = FP32 fat + 16-bit long links")
    ap.add_argument("--delta", type=float, default=0.1,
                    help="reliable update when |r| falls below delta times |r| at the last one")
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-refinement variant")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only launch the hopping terms (FP64, FP32, FP16) and the copy kernel, for counter runs")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: this build has no multi-GPU path "
              "(the lattice would be split by 4-D domain decomposition with halo exchange).",
              file=sys.stderr)
        return 2

    import kuqcd_amd.api as A

    A.set_verbosity(0)
    assert A.kuqcd_init() == A.KUQCD_SUCCESS
    lat = tuple(int(v) for v in args.lattice.split(","))
    Vh = int(np.prod(lat)) // 2

    t = time.time()
    fat, lng = A.gen_gauge(lat, kind=A.KQ_GEN_WEAK, eps=0.3, fat_scale=0.9, c_naik=-1.0 / 24.0, seed_fat=1,
                           seed_long=2)
    b = A.gen_source(lat, 0, seed=3)
    log(f"inputs generated in {time.time() - t:.1f}s")
    A.load_gauge_pair(fat, lng, lat)
    if args.kernels_only:
        for mode, name in ((1, "fp64"), (2, "fp32"), (4, "fp32 fat + 16-bit long"), (3, "fp16")):
            st, r = A.bench_dslash(mode, 1, 0, 0, 20)
            assert st == 0, A.last_error()
            print(f"dslash {name}: {r.ms * 1e3:.1f} us, {r.bytes / Vh:.0f} B/site, {r.gbs:.0f} GB/s, "
                  f"nr {r.nr} nt {r.nt}", flush=True)
        for nr in (2, 4, 8):  # batched FP32 hopping term, links read once for nr right-hand sides
            st, r = A.bench_dslash(2, nr, nr, 256, 20)
            assert st == 0, A.last_error()
            print(f"dslash fp32 x{nr}: {r.ms * 1e3:.1f} us, {r.bytes / Vh:.0f} B/site, {r.gbs:.0f} GB/s", flush=True)
        st, cp = A.bench_copy(1 << 30, 10)
        print(f"copy: 1 GiB read + 1 GiB written per launch, {cp.ms * 1e3:.1f} us, {cp.gbs:.0f} GB/s", flush=True)
        A.kuqcd_gauge_free()
        return 0

    s = np.logspace(np.log10(args.shift_min), 0, args.nshift)
    m2 = args.mass ** 2
    shifts = np.concatenate([[m2 + s[0]], s[1:] - s[0]])  # reference convention (shifts[k>=1] relative)
    b2 = float(np.sum(b * b))
    tol = args.rel_tol ** 2 * b2

    # source and solutions resident in HBM (reference host layout, device memory)
    L = A.L.lib()
    nbytes = b.nbytes
    b_dev = L.kq_device_alloc(nbytes)
    x_dev = [L.kq_device_alloc(nbytes) for _ in range(args.nshift)]
    assert b_dev and all(x_dev)
    assert L.kq_memcpy(b_dev, b.ctypes.data, nbytes) == 0
    import ctypes as C
    xp = (C.c_void_p * args.nshift)(*x_dev)

    def solve(mode, delta=None):
        prm = A.InvertParam(shifts=shifts, tol=tol, max_iter=40000)
        opt = A.default_options(mode=mode, refine=1, rel_tol=args.rel_tol, verbose=0)
        if delta is not None:
            opt.reliable_delta = delta
        rep = A.L.KqSolveReport()
        st = L.kq_solve_multishift_ex(xp, C.c_void_p(b_dev), C.byref(prm.struct()), C.byref(opt), C.byref(rep))
        if st != A.KUQCD_SUCCESS:
            raise RuntimeError(f"solve failed: {st} {A.last_error()}")
        return rep

    for w in range(args.warmup):
        rep = solve(args.mode, args.delta)
        log(f"warmup {w}: {rep.seconds:.3f}s iter {rep.iter}")
    assert L.kq_synchronize() == 0
    t0 = time.perf_counter()
    reps = []
    for k in range(args.steps):
        reps.append(solve(args.mode, args.delta))
        log(f"step {k}: {reps[-1].seconds:.3f}s iter {reps[-1].iter} refine {list(reps[-1].refine_iter[:args.nshift])}")
    assert L.kq_synchronize() == 0
    t1 = time.perf_counter()
    sec_mixed = (t1 - t0) / args.steps
    rep = reps[-1]
    res_mixed = [rep.true_rel_res[k] for k in range(args.nshift)]

    def download():
        out = []
        for p in x_dev:
            h = np.empty_like(b)
            assert L.kq_memcpy(h.ctypes.data, p, nbytes) == 0
            out.append(h)
        return out

    x_mixed = download()

    # the same headline solve with the source and the solutions in host memory (the reference
    # boundary: H2D of b and D2H of the 12 solutions inside the timed region)
    xh = [np.zeros_like(b) for _ in range(args.nshift)]
    xhp = (C.c_void_p * args.nshift)(*[a.ctypes.data for a in xh])

    def solve_host():
        prm = A.InvertParam(shifts=shifts, tol=tol, max_iter=40000)
        opt = A.default_options(mode=args.mode, refine=1, rel_tol=args.rel_tol, verbose=0)
        opt.reliable_delta = args.delta
        rep = A.L.KqSolveReport()
        st = L.kq_solve_multishift_ex(xhp, C.c_void_p(b.ctypes.data), C.byref(prm.struct()), C.byref(opt),
                                      C.byref(rep))
        if st != A.KUQCD_SUCCESS:
            raise RuntimeError(f"solve failed: {st} {A.last_error()}")

    t = time.perf_counter()
    for _ in range(2):
        solve_host()
    sec_host = (time.perf_counter() - t) / 2
    log(f"host-buffer solve: {sec_host:.3f}s")

    # the drop-in entry point itself: kuqcd_solve_multishift on the loaded double pair (the
    # reference method in FP64, its stop test |r_k|^2 < tol with tol = rel^2 |b|^2), host arrays
    prm_ref = A.InvertParam(shifts=shifts, tol=tol, max_iter=40000)
    assert A.kuqcd_solve_multishift(xh, b, prm_ref) == A.KUQCD_SUCCESS, A.last_error()
    t = time.perf_counter()
    for _ in range(2):
        assert A.kuqcd_solve_multishift(xh, b, prm_ref) == A.KUQCD_SUCCESS, A.last_error()
    sec_dropin = (time.perf_counter() - t) / 2
    log(f"drop-in kuqcd_solve_multishift (FP64, host arrays): {sec_dropin:.3f}s iter {prm_ref.iter}")
    del xh

    # tuned variants: the library's default reliable-update interval (delta 0.02), and FP32 fat
    # links with 16-bit fixed-point long links (same FP64 verification of every shift)
    variants = {}
    for key, mode, delta in (("fp32_links_delta_0.02", 2, None), ("fp32_fat_16bit_long_delta_0.02", 4, None)):
        solve(mode, delta)  # warm (autotune / derived formats)
        assert L.kq_synchronize() == 0
        t = time.perf_counter()
        for _ in range(2):
            rv = solve(mode, delta)
        assert L.kq_synchronize() == 0
        sv = (time.perf_counter() - t) / 2
        variants[key] = {"seconds": round(sv, 4), "iterations": rv.iter, "reliable_updates": rv.reliable_updates,
                         "max_true_rel_res": max(rv.true_rel_res[k] for k in range(args.nshift))}
        log(f"variant {key}: {sv:.3f}s iter {rv.iter}")

    # learned refinement in the headline configuration: the network is trained by the library's
    # HIP kernels on solver pairs of 8 seeded 16^4 gauges from the same generator (it is local and
    # gauge covariant, so the weights carry over to 32^4); shift 0 leaves the multishift at 1e-6
    # and the flexible CG finishes it; every shift is then verified in FP64 as usual
    if not args.no_learned:
        lat_t = (16, 16, 16, 16)
        t = time.perf_counter()
        A.learned_clear()
        for sd in range(8):
            f_t, l_t = A.gen_gauge(lat_t, kind=A.KQ_GEN_WEAK, eps=0.3, fat_scale=0.9, c_naik=-1.0 / 24.0,
                                   seed_fat=2 * (100 + sd) + 1, seed_long=2 * (100 + sd) + 2)
            A.load_gauge_pair(f_t, l_t, lat_t)
            A.learned_collect(0, shifts[0], nsrc=1, seed=1000 + sd, rel_tol=1e-11, min_level=1e-8)
        trep = A.learned_train(steps=200, batch=8, lr=3e-4, lr_final=3e-5, seed=1, init=2)
        A.learned_clear()
        t_train = time.perf_counter() - t
        A.load_gauge_pair(fat, lng, lat)

        def solve_learned():
            prm = A.InvertParam(shifts=shifts, tol=tol, max_iter=40000)
            opt = A.default_options(mode=args.mode, refine=1, rel_tol=args.rel_tol, verbose=0, learned=1)
            opt.reliable_delta = args.delta
            rep = A.L.KqSolveReport()
            st = L.kq_solve_multishift_ex(xp, C.c_void_p(b_dev), C.byref(prm.struct()), C.byref(opt), C.byref(rep))
            if st != A.KUQCD_SUCCESS:
                raise RuntimeError(f"learned solve failed: {st} {A.last_error()}")
            return rep

        solve_learned()  # warm
        assert L.kq_synchronize() == 0
        t = time.perf_counter()
        for _ in range(2):
            rl_ = solve_learned()
        assert L.kq_synchronize() == 0
        sv = (time.perf_counter() - t) / 2
        variants["learned_refinement"] = {
            "seconds": round(sv, 4), "iterations": rl_.iter, "reliable_updates": rl_.reliable_updates,
            "shift0_fcg_iterations": rl_.refine_iter[0], "learned_used": rl_.learned_used,
            "max_true_rel_res": max(rl_.true_rel_res[k] for k in range(args.nshift)),
            "training": f"8 gauges 16^4, {trep.nsamples} solver pairs, loss {trep.loss_initial:.3f} -> "
                        f"{trep.loss_final:.3f}, {t_train:.1f}s (outside the timed solves)"}
        log(f"variant learned: {sv:.3f}s iter {rl_.iter} shift-0 FCG {rl_.refine_iter[0]} used {rl_.learned_used}")

    # plain FP64 GPU multishift CG on the same problem
    solve(1)  # warm (autotune / derived formats)
    t2 = time.perf_counter()
    for _ in range(args.fp64_steps):
        rep64 = solve(1)
    assert L.kq_synchronize() == 0
    t3 = time.perf_counter()
    sec_fp64 = (t3 - t2) / args.fp64_steps
    res_fp64 = [rep64.true_rel_res[k] for k in range(args.nshift)]
    # per-shift solution difference, mixed vs the FP64 solve
    x64 = download()
    diff = [float(np.linalg.norm(xm - x6) / np.linalg.norm(x6)) for xm, x6 in zip(x_mixed, x64)]
    log(f"fp64: {sec_fp64:.3f}s iter {rep64.iter}; mixed {sec_mixed:.3f}s")

    # hopping-term roofline (HIP events on the library stream)
    rl = {}
    for mode, name in ((1, "fp64"), (2, "fp32"), (4, "fp32_l16"), (3, "fp16")):
        st, r = A.bench_dslash(mode, 1, 0, 0, 50)
        assert st == 0, A.last_error()
        rl[name] = r
        log(f"dslash {name}: {r.ms * 1e3:.1f} us {r.gbs:.0f} GB/s {100 * r.hbm_frac:.1f}% nt {r.nt}")
    # the mixed solve's dominant kernel: the FP32 term batched over four live systems (the base
    # multishift and forked shifts share it for most iterations), autotuned blocking and cap
    st, r4 = A.bench_dslash(2, 4, 0, 0, 50)
    assert st == 0, A.last_error()
    log(f"dslash fp32 x4: {r4.ms * 1e3:.1f} us {r4.gbs:.0f} GB/s {100 * r4.hbm_frac:.1f}% nr {r4.nr} nt {r4.nt}")
    st, cp = A.bench_copy(1 << 30, 20)
    r64 = rl["fp64"]
    # the FP64 term: long links in 14 reals (LinkD14) when every one is a complex multiple of an
    # SU(3) matrix, else 18 (the algorithmic bytes per site say which ran)
    d14 = round(r64.bytes / Vh) == 2144
    fp64_kernel = (f"k_dslash<LinkD,{'LinkD14' if d14 else 'LinkD'},double> (FP64 hopping term, "
                   f"{r64.bytes / Vh:.0f} B/site)")
    # the committed counter runs are of the 32^4 terms; other lattices carry no traffic figure
    pmc = ROOT / "profiles" / "dslash_32x32x32x32_pmc.txt"
    k64 = "k_dslash<kq::LinkD, kq::LinkD14, double, 1" if d14 else "k_dslash<kq::LinkD, kq::LinkD, double, 1"
    traffic = pmc_traffic(pmc, k64) if lat == (32, 32, 32, 32) else None
    # with 256-thread workgroups on Lx = 32 lattices the batched term stages its x/y neighbour rows in
    # LDS (k_dslash_staged); KQ_STAGE=0 switches that off
    staged_x4 = r4.nt == 256 and lat[0] == 32 and os.environ.get("KQ_STAGE", "1") != "0"
    k_x4 = "k_dslash_staged<kq::LinkF, kq::LinkF12, float, 4" if staged_x4 else "k_dslash<kq::LinkF, kq::LinkF12, float, 4"
    traffic_x4 = pmc_traffic(pmc, k_x4) if lat == (32, 32, 32, 32) else None

    # CPU baseline: the oracle CGM (the reference's algorithm, FP64) solving the same problem to
    # completion on the host cores; its solution also gives each shift's error
    cpu = None
    err_vs_oracle = None
    if not args.no_cpu:
        import oracle
        nthreads = oracle.num_threads()
        bo = np.ascontiguousarray(b)
        t = time.perf_counter()
        xo, it_o = oracle.cgm(fat, lng, lat, 0, shifts, bo, tol, 40000)
        t_cpu = time.perf_counter() - t
        err_vs_oracle = {
            "mixed": [float(f"{np.linalg.norm(a - o) / np.linalg.norm(o):.2e}") for a, o in zip(x_mixed, xo)],
            "fp64": [float(f"{np.linalg.norm(a - o) / np.linalg.norm(o):.2e}") for a, o in zip(x64, xo)],
            "oracle_iterations": it_o,
        }
        cpu = {
            "value": round(t_cpu, 3),
            "unit": "s",
            "cores": nthreads,
            "kind": "port",
            "sample": (f"the whole solve: oracle C restatement of CGM::Invert (FP64, converged shifts dropped as "
                       f"inverter_cgm.cpp:71-76 does), same 32^4 12-shift problem and stop test, {it_o} iterations, "
                       f"{t_cpu / it_o * 1e3:.1f} ms/iteration; {nthreads} OpenMP threads (OMP_PROC_BIND="
                       f"{os.environ.get('OMP_PROC_BIND')}, OMP_PLACES={os.environ.get('OMP_PLACES')}) on "
                       f"{hc['cpus_allowed']} allowed CPUs of a {hc['model']} host with {hc['physical_cores']} "
                       f"physical cores"),
        }
        log(f"cpu: {t_cpu:.1f}s, {it_o} iterations ({nthreads} threads); worst mixed error against the oracle "
            f"{max(err_vs_oracle['mixed']):.2e}")
        # the oracle's FP64 hopping term on the same cores: median of 20 applications
        src_o = np.ascontiguousarray(xo[0])
        ts = []
        for _ in range(20):
            t = time.perf_counter()
            oracle.dslash(fat, lng, src_o, lat, 1)
            ts.append(time.perf_counter() - t)
        t_d = float(np.median(ts))
        cpu["dslash_fp64"] = {"ms": round(t_d * 1e3, 2), "GB/s": round(2400 * Vh / t_d / 1e9, 1),
                              "GFLOP/s": round(1146 * Vh / t_d / 1e9, 1)}

    line = {
        "metric": METRIC,
        "value": round(sec_mixed, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec_mixed * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": MODE_DTYPE[args.mode],
        "data": "synthetic (seeded weak-field SU(3) links, Gaussian source)",
        "config": {"workload": f"HISQ even-odd multishift CG, {'x'.join(map(str, lat))}, {args.nshift} shifts "
                               f"log-spaced {args.shift_min:g}..1, m={args.mass}, FP64 true rel. residual "
                               f"{args.rel_tol:g} per shift",
                   "lattice": list(lat), "nshift": args.nshift, "mass": args.mass,
                   "mode": MODE_NAME[args.mode], "reliable_delta": args.delta, "parallelism": "single GPU",
                   "b_x": "device-resident (reference host layout in HBM); host-buffer time in seconds_host_buffers, "
                          "the drop-in kuqcd_solve_multishift (FP64 reference method) in dropin_fp64_seconds_host_buffers"},
        "seconds_host_buffers": round(sec_host, 4),
        "dropin_fp64_seconds_host_buffers": round(sec_dropin, 4),
        "fp64_seconds": round(sec_fp64, 4),
        "speedup_vs_fp64": round(sec_fp64 / sec_mixed, 3),
        "variants": {k: dict(v, speedup_vs_fp64=round(sec_fp64 / v["seconds"], 3)) for k, v in variants.items()},
        "iterations": {"mixed_multishift": rep.iter, "mixed_reliable_updates": rep.reliable_updates,
                       "mixed_refine": [rep.refine_iter[k] for k in range(args.nshift)], "fp64": rep64.iter},
        "max_true_rel_res": {"mixed": max(res_mixed), "fp64": max(res_fp64)},
        "rel_diff_mixed_vs_fp64": [float(f"{d:.2e}") for d in diff],
        "rel_err_vs_oracle": err_vs_oracle,
        "roofline": {"bound": "hbm", "achieved": round(r64.gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(r64.gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": fp64_kernel, "bytes_per_site": round(r64.bytes / Vh, 1),
                     "us_per_launch": round(r64.ms * 1e3, 2), "gflops": round(r64.gflops, 1)},
        "roofline_fp32": {"achieved": round(rl["fp32"].gbs, 1), "frac": round(rl["fp32"].hbm_frac, 4),
                          "us_per_launch": round(rl["fp32"].ms * 1e3, 2),
                          "bytes_per_site": round(rl["fp32"].bytes / Vh, 1)},
        "roofline_fp32_l16": {"achieved": round(rl["fp32_l16"].gbs, 1), "frac": round(rl["fp32_l16"].hbm_frac, 4),
                              "us_per_launch": round(rl["fp32_l16"].ms * 1e3, 2),
                              "bytes_per_site": round(rl["fp32_l16"].bytes / Vh, 1)},
        "roofline_fp32_x4": {"achieved": round(r4.gbs, 1), "frac": round(r4.hbm_frac, 4),
                             "us_per_launch": round(r4.ms * 1e3, 2), "bytes_per_site": round(r4.bytes / Vh, 1),
                             "nr": r4.nr, "nt": r4.nt, "traffic": traffic_x4,
                             "kernel": (f"{'k_dslash_staged' if staged_x4 else 'k_dslash'}<LinkF,LinkF12,float,4> "
                                        "(the mixed solve's dominant kernel)")},
        "roofline_fp16": {"achieved": round(rl["fp16"].gbs, 1), "frac": round(rl["fp16"].hbm_frac, 4),
                          "us_per_launch": round(rl["fp16"].ms * 1e3, 2),
                          "bytes_per_site": round(rl["fp16"].bytes / Vh, 1)},
        "copy_kernel_gbs": round(cp.gbs, 1),
        "cpu_baseline": cpu,
    }
    for p in [b_dev, *x_dev]:
        L.kq_device_free(p)
    A.kuqcd_gauge_free()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
