"""Autotuner for the stencil launch shapes.

For a grid size it times candidate (bx, by, ich) shapes of the 2.5D stencil
kernels (workgroup tile (bx, by) over (k, j), ``ich`` i-planes marched per
thread) with HIP events around every launch, on the solver's own stream, and
keeps the shape with the smallest device time per step.  Results are cached
in ``tuned_tiles.json`` next to this file, keyed by device architecture and
node count per axis, and applied by ``apply_cached`` (single grids) and
``batch_tiles`` (batches: one (bx, by) for all molecules, ich per grid size).
"""
from __future__ import annotations

import ctypes as C
import json
import os

from . import _lib as L

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_tiles.json")
SHAPES = [(64, 4), (64, 2), (32, 8), (256, 1), (128, 2), (64, 1)]
ICH = [2, 4, 6, 8, 12, 16]
# families whose time the tile choice drives on the fine level
FINE = ("itp", "smooth", "restrict", "prolong", "smooth_itp")


def _arch():
    return os.environ.get("RQ_ARCH", "gfx950")


def load_cache():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save_cache(cache):
    tmp = CACHE + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, CACHE)


def _fine_time(solver, first, steps):
    lib = solver._lib
    nf = len(L.FAMILIES)
    solver._chk(lib.rq_timing_enable(solver._h, 1))
    solver.step(first, steps, poisson_interval=1)
    ms = (C.c_double * nf)()
    cnt = (C.c_int64 * nf)()
    solver._chk(lib.rq_timing_read(solver._h, ms, cnt, nf))
    solver._chk(lib.rq_timing_enable(solver._h, 0))
    return sum(ms[f] for f, name in enumerate(L.FAMILIES) if name in FINE) / steps


def tune_solver(solver, steps=6, shapes=SHAPES, ich=ICH, save=True, log=None):
    """Sweep the fine-level launch shape of an initialised solver; applies and
    returns the best (bx, by, ich) and its ms per step."""
    lib = solver._lib
    step0 = 1 << 20  # far from any schedule boundary of a real run
    best = None
    results = []
    for bx, by in shapes:
        for ic in ich:
            solver._chk(lib.rq_set_tile(solver._h, 0, bx, by, ic))
            solver.step(step0, 2, poisson_interval=1)  # warm the new shape
            step0 += 2
            t = _fine_time(solver, step0, steps)
            step0 += steps
            results.append((t, bx, by, ic))
            if log:
                log(f"[autotune] S={solver.S} {bx}x{by} ich {ic}: {t * 1e3:.1f} us/step")
            if best is None or t < best[0]:
                best = (t, bx, by, ic)
    _, bx, by, ic = best
    solver._chk(lib.rq_set_tile(solver._h, 0, bx, by, ic))
    if save:
        cache = load_cache()
        cache[f"{_arch()}:S{solver.S}"] = {"bx": bx, "by": by, "ich": ic, "ms_per_step": best[0]}
        save_cache(cache)
    return {"bx": bx, "by": by, "ich": ic, "ms_per_step": best[0], "sweep": results}


def apply_cached(solver):
    """Apply a cached fine-level shape for this grid size; True if one existed."""
    ent = load_cache().get(f"{_arch()}:S{solver.S}")
    if not ent:
        return False
    solver._chk(solver._lib.rq_set_tile(solver._h, 0, int(ent["bx"]), int(ent["by"]), int(ent["ich"])))
    return True


def batch_tiles(NNs, default=(64, 4)):
    """(bx, by, {NN: ich}) for a batch from the cache: the most common cached
    (bx, by) among its grid sizes, and each size's cached ich (0 = built-in)."""
    cache = load_cache()
    shapes = {}
    ich = {}
    for nn in set(NNs):
        ent = cache.get(f"{_arch()}:S{nn + 1}")
        if ent:
            shapes[(ent["bx"], ent["by"])] = shapes.get((ent["bx"], ent["by"]), 0) + NNs.count(nn)
            ich[nn] = ent["ich"]
    bx, by = max(shapes, key=shapes.get) if shapes else default
    return bx, by, ich


def tune_sizes(NNs, screen_per_cell=0.1, steps=6, log=None):
    """Tune each grid size on a small synthetic molecule (one H2O at the centre)."""
    from .builders import h2o_single, to_config_atoms
    from .solver import RealQMSolver
    out = {}
    for nn in sorted(set(NNs)):
        levels = 1
        while nn % (1 << levels) == 0 and (nn >> levels) - 1 > 19:
            levels += 1
        levels += 1
        rows, els, screen = h2o_single(nn, screen_per_cell)
        s = RealQMSolver({"gridN": nn, "screen": screen, "atoms": to_config_atoms(rows, els), "poisson": "multigrid",
                          "mg_levels": levels, "bootstrap_poisson": 2})
        s.initialize()
        out[nn] = tune_solver(s, steps=steps, log=log)
        s.close()
    return out
