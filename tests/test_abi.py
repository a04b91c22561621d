"""CPU checks of the C-ABI library: it loads, exports every declared entry
point, its host-side index math and generators are right, and every status
code that the reference returns before touching the device is returned in the
same order (kuqcd.cpp:41-101, 176-199, 303-332, 404-445)."""
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle

ROOT = Path(__file__).resolve().parent.parent


def _exported():
    out = subprocess.run(["nm", "-D", "--defined-only", str(ROOT / "kuqcd_amd/lib/libkuqcd.so")],
                         capture_output=True, text=True, check=True).stdout
    names = [ln.split()[-1] for ln in out.splitlines() if ln.strip()]
    dem = subprocess.run(["c++filt"], input="\n".join(names), capture_output=True, text=True, check=True).stdout
    return set(names), set(dem.splitlines())


def test_exports_every_declared_symbol(kq):
    raw, dem = _exported()
    c_decl = re.findall(r"KQ_API[^;(]*?\b(kq_\w+)\s*\(", (ROOT / "include/kuqcd_hip.h").read_text())
    assert len(c_decl) >= 20
    for name in c_decl:
        assert name in raw, name
        assert hasattr(kq.L.lib(), name)
    cxx_decl = re.findall(r"kuqcd_export[^;(]*?\b(kuqcd_\w+)\s*\(", (ROOT / "include/kuqcd.h").read_text())
    assert sorted(cxx_decl) == sorted(["kuqcd_init", "kuqcd_load_gaugefield", "kuqcd_solve_multishift",
                                       "kuqcd_gauge_free", "kuqcd_end"])
    for name in cxx_decl:
        assert any(d.startswith(f"kuqcd::{name}(") for d in dem), name


def test_soname(kq):
    out = subprocess.run(["readelf", "-d", str(ROOT / "kuqcd_amd/lib/libkuqcd.so")], capture_output=True,
                         text=True, check=True).stdout
    assert "libkuqcd.so.1" in out


def test_struct_layout_matches_reference(kq):
    import ctypes as C
    # kuqcd_enum.h:100-121: 4 unsigned + 7 int-sized fields; 7 ints then double, pointer
    assert C.sizeof(kq.KuQCDGaugeParam) == 44
    assert kq.KuQCDInvertParam.tol.offset == 32 and kq.KuQCDInvertParam.shifts.offset == 40
    assert C.sizeof(kq.KuQCDInvertParam) == 48


@pytest.mark.parametrize("d", [1, 2, 3, 5, 6, 7, 12, 16, 24, 31, 48, 96, 1000, 65535])
def test_fastdiv_exact(kq, d):
    assert kq.L.lib().kq_selftest_fastdiv(d, 1 << 21) == 0


def test_fold_rule_after_divergence(kq):
    """The mixed solver's fold of the low-precision solution updates into the FP64 solutions.
    Normally every shift k < maxTermFold is folded. When the host has flagged the iteration as
    diverged (discard), only the shifts still iterating (k < maxTerm) lose their updates; shifts
    dropped as converged earlier in the same chunk (maxTerm <= k < maxTermFold) keep theirs."""
    f = kq.L.lib().kq_selftest_fold_keeps
    for max_term in range(0, 5):
        for k in range(0, 6):
            assert f(0, k, max_term) == 1
            assert f(1, k, max_term) == (1 if k >= max_term else 0), (k, max_term)


@pytest.mark.parametrize("lat", [(4, 4, 4, 1), (6, 4, 2, 8), (2, 1, 3, 5), (8, 6, 4, 2), (4, 6, 1, 3)])
def test_neighbour_index_matches_oracle(kq, lat):
    Vh = int(np.prod(lat)) // 2
    for q in (0, 1):
        for s in range(Vh):
            for mu in range(4):
                for h in (1, -1, 3, -3):
                    assert kq.neighbour_index(lat, q, s, mu, h) == oracle.nb(lat, q, s, mu, h)


def _cmat(a):
    return (a[..., 0] + 1j * a[..., 1]).reshape(a.shape[0], 4, 3, 3)


@pytest.mark.parametrize("kind", [0, 1])
def test_generator_links(kq, kind):
    lat = (4, 4, 2, 6)
    V = int(np.prod(lat))
    Vh = V // 2
    fat, lng = kq.gen_gauge(lat, kind=kind, eps=0.3, fat_scale=0.9, c_naik=-1 / 24, seed_fat=1, seed_long=2)
    W, _ = kq.gen_gauge(lat, kind=kind, eps=0.3, fat_scale=1.0, seed_fat=2, phases=0, long=False)
    U, N, Wm = _cmat(fat), _cmat(lng), _cmat(W)
    eye = np.eye(3)
    for i in range(V):
        q, s = divmod(i, Vh)
        c = oracle.full_coords(lat, q, s)
        for mu in range(4):
            eta = (-1) ** sum(c[:mu])
            G = U[i, mu] / (0.9 * eta)
            assert np.allclose(G @ G.conj().T, eye, atol=1e-13)
            assert abs(np.linalg.det(G) - 1) < 1e-13
            if kind == 1:
                assert np.linalg.norm(G - eye) < 3.0  # weak field, eps = 0.3
            # long link = c_naik * eta * W(x) W(x+mu) W(x+2mu)
            prod = eye
            for h in range(3):
                cc = list(c)
                cc[mu] = (cc[mu] + h) % lat[mu]
                qq = sum(cc) % 2
                j = qq * Vh + (cc[0] >> 1) + (lat[0] // 2) * (cc[1] + lat[1] * (cc[2] + lat[2] * cc[3]))
                prod = prod @ Wm[j, mu]
            assert np.allclose(N[i, mu], (-1 / 24) * eta * prod, atol=1e-14)


def test_generator_deterministic_and_source(kq):
    lat = (4, 4, 4, 4)
    a, _ = kq.gen_gauge(lat, seed_fat=7, long=False)
    b, _ = kq.gen_gauge(lat, seed_fat=7, long=False)
    c, _ = kq.gen_gauge(lat, seed_fat=8, long=False)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    src = kq.gen_source((16, 16, 16, 16), 0, seed=3)
    z2 = (src ** 2).sum(-1)
    assert abs(z2.mean() - 1.0) < 0.01
    assert not np.array_equal(src, kq.gen_source((16, 16, 16, 16), 0, seed=4))


def test_load_status_codes_before_device(kq):
    g = np.zeros((16 * 4 * 9 * 2,))
    P = kq.gauge_param
    cases = [
        (None, P((2, 2, 2, 2)), kq.KUQCD_ERR_NULL_POINTER),
        (g, None, kq.KUQCD_ERR_NULL_POINTER),
        (g, P((2, 2, 2, 2), order=5), kq.KUQCD_ERR_INVALID_PARAM),
        (g, P((2, 2, 2, 2), link_type=kq.KUQCD_WILSON_LINK), kq.KUQCD_ERR_INVALID_PARAM),
        (g, P((2, 2, 2, 2), recon=kq.KUQCD_GAUGE_RECON_12), kq.KUQCD_ERR_UNSUPPORTED),
        (g, P((2, 2, 2, 2), recon=kq.KUQCD_GAUGE_RECON_8), kq.KUQCD_ERR_UNSUPPORTED),
        (g, P((2, 2, 2, 2), recon=5), kq.KUQCD_ERR_INVALID_PARAM),
        (g, P((2, 0, 2, 2)), kq.KUQCD_ERR_INVALID_PARAM),
        (g, P((2, 2, 2, 2), prec=kq.KUQCD_HALF), kq.KUQCD_ERR_UNSUPPORTED),
        (g, P((2, 2, 2, 2), prec=3), kq.KUQCD_ERR_INVALID_PARAM),
        (g, P((2, 2, 2, 2), site_subset=7), kq.KUQCD_ERR_INVALID_PARAM),
        (g, P((2, 2, 2, 2), halo_depth=2), kq.KUQCD_ERR_INVALID_PARAM),
        (g, P((2, 2, 2, 2), site_subset=kq.KUQCD_EVEN), kq.KUQCD_ERR_INVALID_PARAM),
        (g, P((3, 2, 2, 2)), kq.KUQCD_ERR_INVALID_PARAM),
    ]
    kq.set_verbosity(0)
    for arr, prm, want in cases:
        assert kq.kuqcd_load_gaugefield(arr, prm) == want, (prm and list(prm.lat), want)
    # reference check order: null before order before link type before recon before lat before prec
    assert kq.kuqcd_load_gaugefield(g, P((0, 2, 2, 2), order=5, prec=kq.KUQCD_HALF)) == kq.KUQCD_ERR_INVALID_PARAM
    assert kq.kuqcd_load_gaugefield(g, P((0, 2, 2, 2), recon=12, prec=3)) == kq.KUQCD_ERR_UNSUPPORTED


def test_solve_status_codes_before_device(kq):
    kq.set_verbosity(0)
    b = np.zeros((8, 3, 2))
    x = [np.zeros_like(b)]
    prm = kq.InvertParam(shifts=np.array([0.1]), tol=1e-10)
    assert kq.kuqcd_solve_multishift(None, b, prm) == kq.KUQCD_ERR_NULL_POINTER
    assert kq.kuqcd_solve_multishift(x, None, prm) == kq.KUQCD_ERR_NULL_POINTER
    assert kq.kuqcd_solve_multishift(x, b, None) == kq.KUQCD_ERR_NULL_POINTER
    prm.matpc_type = 3
    assert kq.kuqcd_solve_multishift(x, b, prm) == kq.KUQCD_ERR_INVALID_PARAM
    prm.matpc_type = kq.KUQCD_EVEN_EVEN
    # nothing loaded
    assert kq.kuqcd_solve_multishift(x, b, prm) == kq.KUQCD_ERR_UNKNOWN_TYPE


def test_init_grid(kq):
    kq.set_verbosity(0)
    assert kq.kuqcd_init((1, 1, 1, 1)) == kq.KUQCD_SUCCESS
    assert kq.kuqcd_init((2, 1, 1, 1)) == kq.KUQCD_ERR_INVALID_PARAM


@pytest.mark.parametrize("text,want", [("gridx=1\ngridy=1\ngridz=1\ngridt=1\n", 0),
                                       ("gridx = 1\nbogus\nfoo=3\ngridt=1\n", 0),
                                       ("gridx=2\ngridy=1\ngridz=1\ngridt=1\n", -1)])
def test_init_reads_stdin(kq, text, want):
    """kuqcd_init parses gridx..gridt from stdin (command_line.cpp:42-72)."""
    code = ("import sys; sys.path.insert(0, %r); import kuqcd_amd.api as A; "
            "sys.exit(A.kuqcd_init_stdin() & 0xff)" % str(ROOT))
    r = subprocess.run([sys.executable, "-c", code], input=text, capture_output=True, text=True, timeout=120)
    assert r.returncode == (want & 0xff), r.stderr


def test_cpp_dropin_compiles():
    """A C++ caller of the reference API (include/kuqcd.h) builds and links against libkuqcd.so."""
    out = ROOT / "tests" / "cpp" / "dropin_test_solver"
    r = subprocess.run(["make", "-s", "-C", str(ROOT / "tests" / "cpp")], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert out.exists()


INNER = [("float", 1), ("double", 1), ("float", 16), ("double", 8)]


@pytest.mark.parametrize("T,nb", INNER)
def test_exports_inner_classes(kq, T, nb):
    """kuqcd_classes.h's instantiations (dirac_hisq.h:60-68, spinor_field.h:76-84, inverter_cgm.h)
    are exported with default visibility."""
    _, dem = _exported()
    sf = f"kuqcd::SpinorField<{T}, {nb}, 1>"
    gf = f"kuqcd::GaugeField<{T}, {nb}, (KuQCDReconstructType)18>"
    dh = f"kuqcd::DiracHISQ<{T}, {nb}, (KuQCDReconstructType)18, (KuQCDReconstructType)18>"
    want = [f"{sf}::SpinorField(kuqcd::LattBase&)", f"{sf}::SetValues({T} const*)", f"{sf}::GetValues({T}*) const",
            f"{sf}::Clear()", f"{gf}::GaugeField(kuqcd::LattBase&, kuqcd::CommBase&)", f"{gf}::SetValues({T} const*)",
            f"{dh}::Dslash({sf} const&, {sf}&, bool)", f"{dh}::MdaggM({sf} const&, {sf}&, bool)",
            f"kuqcd::CGM<{T}, {sf} >::Invert(int, {T}*, {sf} const&, {sf}**, double, int)",
            f"{T} kuqcd::Norm2<{T}, {nb}, 1>({sf} const&)", f"{T} kuqcd::RealDot<{T}, {nb}, 1>({sf} const&, {sf} const&)",
            f"void kuqcd::Axpby<{T}, {nb}, 1>({sf}&, {sf} const&, {T}, {T})"]
    for w in want:
        assert w in dem, w


def test_cpp_callers_build_without_a_gpu():
    """The C++ callers compile against include/kuqcd.h alone and link against libkuqcd.so."""
    subprocess.run(["make", "-s", "-C", str(ROOT / "tests" / "cpp")], check=True, timeout=300)
    assert (ROOT / "tests" / "cpp" / "inner_api_test").exists()
