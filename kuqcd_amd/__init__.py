"""MI355X implementation of the KuQCD HISQ even-odd multishift CG solver.

The compute path is libkuqcd.so (hand-written HIP kernels for gfx950 plus a
C++ host runtime); this package is its Python binding. See include/kuqcd.h
(the drop-in C++ API) and include/kuqcd_hip.h (the C ABI).
"""
from . import api  # noqa: F401
from ._lib import LIB_PATH, last_error, load  # noqa: F401

__all__ = ["api", "load", "last_error", "LIB_PATH"]
