"""Adjoint forces at full size against a noise-free reference.

Central differences of the FP32 solve cannot resolve 1e-4 Ha/bohr at these
sizes: the rounding of a -84 Ha energy (1 ulp ~ 8e-6 Ha) over 2 delta =
2e-3 bohr is already ~4e-3 Ha/bohr.  The derivative of the same map in FP64
is what those differences tend to in exact arithmetic, so the GPU's reverse
sweep is compared with autograd of the FP64 oracle of the map
(oracle/adjoint.py, run on the GPU in FP64 to make the full grid
affordable): the 10-water 192^3 workload with its 5-level V-cycle (20-step
map) and the tri-glycine + 9-water 256^3 workload (51 atoms, 10-step map),
each from a start converged at the reference geometry (the library's
saved-start map), at a geometry displaced by a seeded 0.02 bohr jitter.
Bars: every force component within 1e-4 Ha/bohr (cloud/solver.py:282-315 is
the reference's own force; the adjoint differentiates the energy the
reference reports, cloud/solver.py:227-280).  The energy's difference is the
FP32 rounding of the map against FP64 and is printed: 2.5e-7 relative at
192^3; 4.4e-6 at 256^3, where K is a 51-term FP32 sum over atoms of Z up to
4 and V_eK is about twice the total, so the bar there is 1e-5.
"""
import numpy as np
import pytest

from oracle import adjoint as OA
from oracle import s1
from realmolecule_amd import builders as B
from realmolecule_amd.adjoint import FrozenSolver

pytestmark = pytest.mark.gpu


def _compare(w, n_steps, n_converge, jitter, seed):
    NN, screen, rows = w["NN"], w["screen"], w["rows"]
    print(f"[fullsize] NN={NN} atoms={len(rows)} levels={w['mg_levels']} steps={n_steps}", flush=True)
    fz = FrozenSolver(rows, NN, screen, n_steps, bootstrap=10, checkpoint=0, poisson_interval=1, norm_interval=20,
                      poisson="multigrid", mg_levels=w["mg_levels"])
    pos = rows[:, :3] * fz.h
    fz.warm_start(pos, n_converge)
    U0, P0 = fz.s.U.get().astype(np.float64), fz.s.P.get().astype(np.float64)
    lab = fz.s.label.get()
    pos2 = pos + np.random.default_rng(seed).normal(0.0, jitter, pos.shape)
    E, g = fz.energy_and_gradient(pos2)
    dv = fz.s.dv
    fz.close()
    rows2 = rows.copy()
    rows2[:, :3] = (pos2 / (screen / NN)).astype(np.float32)
    h = screen / NN
    Eo, go = OA.energy_and_gradient(rows2, lab, NN, screen, n_steps, poisson_interval=1, norm_interval=20,
                                    start=(U0, P0), mode="multigrid", mg_levels=w["mg_levels"], dv=dv,
                                    device="cuda")
    go = go - s1.nuclear_forces(rows2[:, :3], rows2[:, 3], h)
    Eo += s1.nuclear_energy(rows2[:, :3], rows2[:, 3], h)
    err = np.abs(g - go)
    print(f"[fullsize] E gpu {E:.9f} fp64 {Eo:.9f} (rel {abs(E - Eo) / abs(Eo):.1e}); max |F| {np.abs(go).max():.3f}, "
          f"max |dF| {err.max():.2e} Ha/bohr (atom {np.unravel_index(err.argmax(), err.shape)})", flush=True)
    return E, Eo, err


def test_c2_adjoint_forces_match_fp64_autograd_at_full_size():
    E, Eo, err = _compare(B.c2_waters(), n_steps=20, n_converge=1000, jitter=0.02, seed=11)
    assert abs(E - Eo) <= 1e-6 * abs(Eo)
    assert err.max() <= 1e-4


def test_c4_adjoint_forces_match_fp64_autograd_at_full_size():
    E, Eo, err = _compare(B.c4_peptide(), n_steps=10, n_converge=1000, jitter=0.02, seed=12)
    assert abs(E - Eo) <= 1e-5 * abs(Eo)
    assert err.max() <= 1e-4
