"""The inner classes of kuqcd.h (dirac.h:25-26, dirac_hisq.h:57-58, inverter_cgm.h:25,
spinor_field.h:42-64), called from C++ (tests/cpp/inner_api_test.cpp, compiled with g++
against the header) on fields in GPU memory, against the oracle:

  * DiracHISQ::Dslash (odd -> even) and MdaggM: <= 1e-13 relative in double, <= 1e-5 in float
    (the oracle's float operator rounds in another order).
  * Norm2, RealDot, Axpby against numpy on the downloaded fields.
  * CGM::Invert on the library's DiracHISQ (the GPU runner) and on a caller's own Dirac
    subclass (the host-driven recurrence): iteration counts within 1 % (at least 2) of the
    oracle's CGM and every shift's solution within 1e-8 (double; float: 5e-6, where each float
    solution lies ~7e-7 from the FP64 one).
  * KUQCD_HISQ_DIRAC on full-lattice fields: each parity of the result is the even-odd operator
    applied to the other parity of the source.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from conftest import rel_err

HERE = Path(__file__).resolve().parent


def _build():
    subprocess.run(["make", "-s", "-C", str(HERE / "cpp")], check=True)
    return HERE / "cpp" / "inner_api_test"


@pytest.mark.gpu
@pytest.mark.parametrize("mode,lat", [("double", (8, 8, 8, 8)), ("float16", (32, 4, 4, 8))])
def test_inner_classes_against_oracle(tmp_path, kq, mode, lat):
    exe = _build()
    dt = np.float64 if mode == "double" else np.float32
    fat, lng = kq.gen_gauge(lat, kind=kq.KQ_GEN_HAAR, fat_scale=0.9, seed_fat=1, seed_long=2)
    fat, lng = fat.astype(dt), lng.astype(dt)
    b = kq.gen_source(lat, 0, seed=3).astype(dt)
    src = kq.gen_source(lat, 1, seed=4).astype(dt)
    V = int(np.prod(lat))
    full = np.concatenate([kq.gen_source(lat, 0, seed=5), kq.gen_source(lat, 1, seed=6)]).astype(dt)
    m = 0.05
    s = np.array([0.0, 0.01, 0.1, 1.0])
    shifts = np.concatenate([[m * m + s[0]], s[1:] - s[0]])
    rel = 1e-10 if mode == "double" else 1e-5
    tol = rel * rel * float(np.sum(b.astype(np.float64) ** 2))
    max_iter = 5000
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        np.array(lat, np.int32).tofile(f)
        np.array([len(shifts), max_iter], np.int32).tofile(f)
        np.array([tol], np.float64).tofile(f)
        shifts.astype(np.float64).tofile(f)
        for a in (fat, lng, b, src, full):
            np.ascontiguousarray(a).tofile(f)
    out = tmp_path / "out.bin"
    print("inner API caller:", mode, lat, flush=True)
    r = subprocess.run([str(exe), mode, str(inp), str(out)], capture_output=True, text=True, timeout=120,
                       input="gridx=1\ngridy=1\ngridz=1\ngridt=1\n")
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr

    n = b.size
    raw = out.read_bytes()
    pos = 0

    def take(count, typ):
        nonlocal pos
        a = np.frombuffer(raw, dtype=typ, count=count, offset=pos)
        pos += a.nbytes
        return a

    dslash = take(n, dt).reshape(b.shape)
    mdm = take(n, dt).reshape(b.shape)
    norm2, dot = take(2, np.float64)
    axpby = take(n, dt).reshape(b.shape)
    it = int(take(1, np.int32)[0])
    sol = [take(n, dt).reshape(b.shape) for _ in shifts]
    it2 = int(take(1, np.int32)[0])
    sol2 = [take(n, dt).reshape(b.shape) for _ in shifts]
    fulld = take(full.size, dt).reshape(full.shape)
    assert pos == len(raw)

    op_tol = 1e-13 if mode == "double" else 1e-5
    assert rel_err(dslash, oracle.dslash(fat, lng, src, lat, 0)) <= op_tol
    ref_mdm = oracle.mdaggm(fat, lng, b, lat, 0)
    assert rel_err(mdm, ref_mdm) <= op_tol
    b64, m64 = b.astype(np.float64), mdm.astype(np.float64)
    # the sums are formed in FP64 and returned as T (spinor_field.h:48-64)
    red_tol = 1e-12 if mode == "double" else 1e-6
    assert abs(norm2 - np.sum(b64 * b64)) <= red_tol * np.sum(b64 * b64)
    assert abs(dot - np.sum(b64 * m64)) <= red_tol * np.linalg.norm(b64) * np.linalg.norm(m64)
    assert rel_err(axpby, 0.75 * b64 - 1.25 * m64) <= (1e-15 if mode == "double" else 1e-6)

    xo, ito = oracle.cgm(fat, lng, lat, 0, shifts.astype(dt), b, tol, max_iter)
    print(f"CGM iterations: device runner {it}, host-driven {it2}, oracle {ito}", flush=True)
    sol_tol = 1e-8 if mode == "double" else 5e-6
    for iters, xs in ((it, sol), (it2, sol2)):
        assert abs(iters - ito) <= max(2, ito // 100), (iters, ito)
        for k in range(len(shifts)):
            assert rel_err(xs[k], xo[k]) <= sol_tol, k

    Vh = V // 2
    assert rel_err(fulld[:Vh], oracle.dslash(fat, lng, np.ascontiguousarray(full[Vh:]), lat, 0)) <= op_tol
    assert rel_err(fulld[Vh:], oracle.dslash(fat, lng, np.ascontiguousarray(full[:Vh]), lat, 1)) <= op_tol
