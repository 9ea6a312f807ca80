"""realmolecule_amd: MI355X-native (gfx950 HIP) RealQM grid solver.

The hot path -- label-Neumann imaginary-time update, Poisson (Jacobi /
multigrid V-cycle), per-domain normalisation, domain-boundary rules, FP64
energy and force reductions -- runs in ``librq.so`` behind the C ABI of
``include/rq.h``.  ``RealQMSolver`` mirrors ``cloud/solver.py``'s class.
"""
from ._lib import load as load_library, LIB_PATH  # noqa: F401


def __getattr__(name):
    if name == "RealQMSolver":
        from .solver import RealQMSolver
        return RealQMSolver
    raise AttributeError(name)
