"""Build libkuqcd.so (HIP kernels + host C++ runtime) for gfx950, in-tree.

Every translation unit is compiled by hipcc for gfx950 with packed-FP32
instructions disabled, in parallel, and linked into
kuqcd_amd/lib/libkuqcd.so (SONAME libkuqcd.so.1, as the reference's
src/CMakeLists.txt:33, 43). The oracle (test infrastructure) is built by
oracle/Makefile.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
OBJDIR = PKG / "_build"
LIB = LIBDIR / "libkuqcd.so"

SOURCES = [
    "kq_dslash_fp64.hip",
    "kq_dslash_fp64r.hip",
    "kq_dslash_fp32.hip",
    "kq_dslash_mixed.hip",
    "kq_dslash_16.hip",
    "kq_dslash_u3.hip",
    "kq_cgm_kernels.hip",
    "kq_layout.hip",
    "kq_ptnet.hip",
    "kq_fcg.hip",
    "kq_blas.hip",
    "kq_noalloc.hip",
    "kq_api.cpp",
    "kq_solve.cpp",
    "kq_learned.cpp",
    "kq_ops.cpp",
    "kq_gen.cpp",
    "kq_classes.cpp",
    "kq_xfer.cpp",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-fvisibility=hidden",
    "-Wall",
    "-Wno-unused-result",
    # no packed FP32 (v_pk_*_f32) instructions anywhere
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


def _compile(src: str, verbose: bool) -> Path:
    obj = OBJDIR / (Path(src).stem + ".o")
    cmd = [HIPCC, *COMMON, "-x", "hip", "-c", str(CSRC / src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile every source and link libkuqcd.so; returns its path."""
    OBJDIR.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    missing = [s for s in SOURCES if not (CSRC / s).exists()]
    if missing:
        raise FileNotFoundError(f"missing sources: {missing}")
    srcs = list(SOURCES)
    jobs = jobs or min(len(srcs), max(1, min(16, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp),
           "-Wl,-soname,libkuqcd.so.1", "-Wl,-z,defs", "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    so1 = LIBDIR / "libkuqcd.so.1"
    try:
        if so1.is_symlink() or so1.exists():
            so1.unlink()
        so1.symlink_to("libkuqcd.so")
    except OSError:
        shutil.copyfile(LIB, so1)
    return LIB


def build_oracle(verbose: bool = False) -> None:
    """Build the oracle restatement (test infrastructure)."""
    cmd = ["make", "-s", "-C", str(ROOT / "oracle")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if verbose or r.returncode != 0:
        print(r.stdout, r.stderr, file=sys.stderr)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    build_oracle()
