"""ctypes face of ``libcstep.so`` (oracle/cstep.c): the C + OpenMP
restatement of the cloud/solver.py step with a multigrid Poisson step
(TEST INFRASTRUCTURE ONLY -- the all-core CPU baseline of bench.py and a
check in tests/test_oracle_cstep.py against the NumPy oracle).

Pinned by tests/test_oracle_cstep.py: the same steps as oracle/s1.py passes
with oracle/mg.py's vcycle on the same inputs, to a few FP32 ulps (the
per-domain FP64 norm sums and the FP64 coarsest solve add in a different
order).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def load():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "libcstep.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: build it with `make -C oracle`")
        lib = C.CDLL(path)
        f = C.c_float
        P = C.c_void_p
        lib.cs_steps.restype = C.c_int
        lib.cs_steps.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, f, f, f, f, f, f, f, P, P, P, P, P, C.c_int,
                                 C.c_long, C.c_int, C.c_int, C.c_int]
        lib.cs_threads.restype = C.c_int
        _lib = lib
    return _lib


def threads() -> int:
    return int(load().cs_threads())


def run_steps(U, P, K, label, Z, g, levels, first, n, nu=2, a=1.0 / 7.0, c=7.0, poisson_interval=1,
              norm_interval=20):
    """Steps first+1..first+n in place on float32 (S,S,S) U and P; ``g`` is
    an oracle.s1.Grid in float32 (its FP32 constants are used)."""
    lib = load()
    for arr, dt in ((U, np.float32), (P, np.float32), (K, np.float32), (label, np.int32)):
        assert arr.dtype == dt and arr.flags.c_contiguous
    Zf = np.ascontiguousarray(Z, np.float32)
    rc = lib.cs_steps(g.NN, levels, nu, nu, float(np.float32(a)), float(np.float32(c)), float(g.h2), float(g.h3),
                      float(g.two_pi), float(g.half_dv), float(g.dtv), U.ctypes.data, P.ctypes.data, K.ctypes.data,
                      label.ctypes.data, Zf.ctypes.data, len(Zf), int(first), int(n), int(poisson_interval),
                      int(norm_interval))
    if rc != 0:
        raise ValueError("cs_steps: bad arguments or allocation failure")
