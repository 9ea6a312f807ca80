"""Differentiable energy: dE/dR of the fixed-step solve by the hand-written
adjoint in librq.so, behind a ``torch.autograd.Function``.

The differentiated map (frozen labels, see ``rq_adjoint_gradient`` in
include/rq.h) replaces the plain path's forces by central finite
differences of full solves (6N+1 runs of ``RealQMSolver.run``,
``cloud/solver.py:550-634``): one forward solve plus one reverse sweep with
checkpointed recomputation of the time-stepping trajectory.

Units: positions in bohr (absolute, box origin at node 0), energies in Ha,
forces in Ha/bohr.  The nucleus-nucleus energy E_KK (h^2-softened,
cloud/solver.py:462-474) and its gradient are added on the host in FP64.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L
from .solver import RealQMSolver, atoms_array, nuclear_energy, nuclear_forces


def make_spec(n_steps, bootstrap=100, checkpoint=0, poisson_interval=2, norm_interval=20):
    """rq_adjoint_spec; ``checkpoint`` <= 0 lets the library keep the whole
    trajectory when it fits (no recomputation in the reverse sweep)."""
    sp = L.RqAdjointSpec()
    sp.n_steps = int(n_steps)
    sp.bootstrap = int(bootstrap)
    sp.checkpoint_interval = int(checkpoint)
    sp.schedule = L.RqSchedule(int(poisson_interval), int(norm_interval), 1 << 30)
    return sp


class FrozenSolver:
    """A solver handle whose labels are frozen at a reference geometry.

    ``rows`` are (i, j, k, Z, rc) in grid units; the labels come from the
    trial argmax at that geometry (cloud/solver.py:36-53)."""

    def __init__(self, rows, NN, screen, n_steps, bootstrap=100, checkpoint=0, poisson_interval=2,
                 norm_interval=20, **cfg):
        from .builders import to_config_atoms
        self.rows = np.asarray(rows, np.float64).copy()
        self.NN, self.screen = NN, screen
        self.h = screen / NN
        config = {"gridN": NN, "screen": screen, "atoms": to_config_atoms(self.rows), "boundary": "off"}
        config.update(cfg)
        self.s = RealQMSolver(config)
        self.lib, self.hdl = self.s._lib, self.s._h
        self.s._chk(self.lib.rq_init_trial(self.hdl))  # reference labels
        self.spec = make_spec(n_steps, bootstrap, checkpoint, poisson_interval, norm_interval)
        self.n = len(self.rows)

    def _set_positions(self, pos_bohr):
        rows = self.rows.copy()
        rows[:, :3] = np.asarray(pos_bohr, np.float64) / self.h
        self.s._chk(self.lib.rq_set_atoms(self.hdl, atoms_array(rows), self.n))
        return rows

    def energy(self, pos_bohr):
        """Total energy (electronic + E_KK) of the frozen-label solve."""
        rows = self._set_positions(pos_bohr)
        e3 = (C.c_double * 3)()
        self.s._chk(self.lib.rq_solve_frozen(self.hdl, C.byref(self.spec), e3))
        return float(e3[0] + e3[1] + e3[2]) + self._ekk(rows)

    def energy_and_gradient(self, pos_bohr):
        """(E, dE/dR[n,3] per bohr) by one forward solve + the reverse sweep."""
        rows = self._set_positions(pos_bohr)
        e3 = (C.c_double * 3)()
        g = (C.c_double * (3 * self.n))()
        self.s._chk(self.lib.rq_adjoint_gradient(self.hdl, C.byref(self.spec), e3, g))
        grad = np.frombuffer(g, np.float64).reshape(self.n, 3).copy()
        grad -= nuclear_forces(rows[:, :3].astype(np.float32), rows[:, 3], self.h)
        return float(e3[0] + e3[1] + e3[2]) + self._ekk(rows), grad

    def _ekk(self, rows):
        return nuclear_energy(rows[:, :3].astype(np.float32), rows[:, 3], self.h)

    def warm_start(self, pos_bohr, n_converge, bootstrap=None):
        """Converge at ``pos_bohr`` (the cold map with ``n_converge`` steps),
        store that U and P as the start of every later solve
        (RQ_ADJ_START_SAVED) and return its energy.  Later energies and
        gradients are those of the spec's n_steps from the stored state, at
        any positions; the stored state does not move with them."""
        rows = self._set_positions(pos_bohr)
        cold = make_spec(n_converge, self.spec.bootstrap if bootstrap is None else bootstrap,
                         1, self.spec.schedule.poisson_interval, self.spec.schedule.norm_interval)
        e3 = (C.c_double * 3)()
        self.s._chk(self.lib.rq_solve_frozen(self.hdl, C.byref(cold), e3))
        self.s._chk(self.lib.rq_adjoint_save_start(self.hdl))
        self.spec.start = L.RQ_ADJ_START_SAVED
        return float(e3[0] + e3[1] + e3[2]) + self._ekk(rows)

    def refreeze(self, pos_bohr):
        """Re-freeze the labels at new nuclei: trial argmax at ``pos_bohr``."""
        self._set_positions(pos_bohr)
        self.s._chk(self.lib.rq_init_trial(self.hdl))

    def stored_position(self, x_bohr):
        """The position the solver actually uses: grid units held as FP32."""
        return float(np.float32(x_bohr / self.h)) * self.h

    def forces_fd(self, pos_bohr, delta=1e-3, atoms=None):
        """Plain path: central differences of full solves (2 per component).

        The step is divided by the difference of the stored (FP32 grid-unit)
        positions, not by 2 delta: at |R| ~ 20 bohr an FP32 grid coordinate
        moves in steps of ~2e-6 bohr, a 1e-4 relative error on a 1e-2 bohr
        step."""
        pos = np.asarray(pos_bohr, np.float64)
        idx = range(self.n) if atoms is None else atoms
        F = np.zeros((self.n, 3))
        for a in idx:
            for c in range(3):
                p = pos.copy()
                p[a, c] = pos[a, c] + delta
                xp = self.stored_position(p[a, c])
                ep = self.energy(p)
                p[a, c] = pos[a, c] - delta
                xm = self.stored_position(p[a, c])
                em = self.energy(p)
                F[a, c] = -(ep - em) / (xp - xm)
        return F

    def close(self):
        self.s.close()


class TrajectoryEnergy(torch.autograd.Function):
    """E(R) of the frozen-label solve; backward returns dE/dR from the adjoint."""

    @staticmethod
    def forward(ctx, positions, frozen: FrozenSolver):
        pos = positions.detach().cpu().double().numpy()
        E, grad = frozen.energy_and_gradient(pos)
        ctx.save_for_backward(torch.as_tensor(grad, dtype=positions.dtype, device=positions.device))
        return positions.new_tensor(E)

    @staticmethod
    def backward(ctx, gE):
        (grad,) = ctx.saved_tensors
        return gE * grad, None


def trajectory_energy(positions: torch.Tensor, frozen: FrozenSolver) -> torch.Tensor:
    return TrajectoryEnergy.apply(positions, frozen)
