"""Geometry relaxation on adjoint forces (L-BFGS), with the finite-difference
baseline.

The reference moves nuclei by damped Euler steps on Hellmann-Feynman forces
(``cloud/solver.py:522-537`` ``_move_nuclei``; ``molecule.js:4410-4524``);
``RealQMSolver._move_nuclei`` keeps that integrator.  This driver instead
minimises the frozen-label trajectory energy E(R) of ``FrozenSolver``
(forward solve + hand-written adjoint, ``rq_adjoint_gradient``) with
scipy's L-BFGS-B.  Labels are re-frozen (trial argmax at the new nuclei,
``cloud/solver.py:36-53``) after every accepted iteration, so each line search
sees one smooth E(R).  ``fd=True`` replaces the adjoint gradient by central
differences of full solves: the plain path, 6N+1 solves per gradient.
Positions in bohr, energies in Ha.
"""
from __future__ import annotations

import time

import numpy as np
from scipy.optimize import minimize

from .adjoint import FrozenSolver


def relax(rows, NN, screen, n_steps=200, bootstrap=20, max_iter=20, gtol=1e-3, fd=False, fd_delta=1e-2,
          log=None, converge=0, ftol=None, refreeze=None, **cfg):
    """Relax the nuclei of ``rows`` ((i, j, k, Z, rc) in grid units).

    ``converge`` > 0: the energy is that of the ``n_steps`` that follow a
    state converged by ``converge`` steps at the start geometry
    (FrozenSolver.warm_start), one fixed function of R for the whole run
    (labels are then not re-frozen unless ``refreeze`` is True).
    ``ftol``: scipy's relative-reduction stop (None: its default).

    Returns a dict: final rows, energy history (one entry per energy
    evaluation), gradient norms per accepted iteration, solver evaluations,
    wall time and scipy's status message."""
    fz = FrozenSolver(rows, NN, screen, n_steps, bootstrap=bootstrap, **cfg)
    x0 = np.asarray(rows, np.float64)[:, :3] * fz.h
    if refreeze is None:
        refreeze = converge <= 0
    t_start = time.perf_counter()
    if converge > 0:
        fz.warm_start(x0, converge)
    t_start = time.perf_counter() - t_start
    energies, gnorms = [], []
    last = {}

    def fun(x):
        pos = x.reshape(-1, 3)
        if fd:
            E = fz.energy(pos)
            g = -fz.forces_fd(pos, delta=fd_delta)
        else:
            E, g = fz.energy_and_gradient(pos)
        energies.append(E)
        last["g"] = g
        if log and fd:
            log(f"[relax] central-difference evaluation {len(energies)}: E {E:.8f}")
        return E, g.ravel()

    def accepted(xk):
        gnorms.append(float(np.abs(last["g"]).max()))
        if log:
            log(f"[relax] iter {len(gnorms)}: E {energies[-1]:.8f} max|dE/dR| {gnorms[-1]:.3e}")
        if refreeze:
            fz.refreeze(xk.reshape(-1, 3))

    t0 = time.perf_counter()
    # nuclei stay 5 cells inside the box, the reference clip (cloud/solver.py:536)
    lo, hi = 5 * fz.h, (NN - 5) * fz.h
    opts = {"maxiter": int(max_iter), "gtol": float(gtol)}
    if ftol is not None:
        opts["ftol"] = float(ftol)
    res = minimize(fun, x0.ravel(), jac=True, method="L-BFGS-B", callback=accepted,
                   bounds=[(lo, hi)] * x0.size, options=opts)
    wall = time.perf_counter() - t0
    out_rows = np.asarray(rows, np.float64).copy()
    out_rows[:, :3] = res.x.reshape(-1, 3) / fz.h
    fz.close()
    return {"rows": out_rows, "energies": energies, "grad_max": gnorms, "evaluations": len(energies),
            "wall_s": wall, "start_s": t_start, "message": str(res.message), "success": bool(res.success)}
