"""Batched reverse sweep: the frozen-label map and its adjoint over several
molecules with uneven grids in one launch per pass (rq_batch_adjoint_gradient)
against each molecule's own handle (rq_adjoint_gradient through
FrozenSolver), at the reference geometry and after moving one molecule's
atoms.  The passes run the same column bodies, so energies and gradients
agree to FP64 rounding of the per-thread partial sums."""
import numpy as np
import pytest

from realmolecule_amd.adjoint import FrozenSolver, make_spec
from realmolecule_amd.batch import Batch
from realmolecule_amd.solver import nuclear_forces
from tests.helpers import three_atoms

pytestmark = pytest.mark.gpu


def _mols():
    out = []
    for NN in (32, 40, 48):
        rows = three_atoms(NN)
        rows[:, :3] += np.array([0.13, -0.21, 0.07]) * (NN / 32)
        out.append((rows, NN, NN * 0.3))
    return out


def _single(rows, NN, screen, n_steps, pos_rows=None, checkpoint=8):
    fz = FrozenSolver(rows, NN, screen, n_steps, bootstrap=4, checkpoint=checkpoint, poisson_interval=1,
                      poisson="multigrid", mg_levels=3)
    r = rows if pos_rows is None else pos_rows
    E, g = fz.energy_and_gradient(r[:, :3] * fz.h)
    g_el = g + nuclear_forces(r[:, :3].astype(np.float32), r[:, 3], fz.h)
    E_el = E - fz._ekk(r)
    fz.close()
    return E_el, g_el


@pytest.mark.parametrize("checkpoint", [8, 64])
def test_batched_adjoint_matches_single_handles(checkpoint):
    mols = _mols()
    n_steps = 30
    b = Batch(mols, mg_levels=3, graph=False)
    b.initialize(bootstrap=0)
    spec = make_spec(n_steps, bootstrap=4, checkpoint=checkpoint, poisson_interval=1, norm_interval=20)
    print(f"batched adjoint, checkpoint {checkpoint}", flush=True)
    E3, grads = b.adjoint_gradient(spec)
    # a moved molecule: the frozen map at new positions, labels kept
    moved = mols[1][0].copy()
    moved[:, :3] += np.array([[0.3, -0.2, 0.1], [0.0, 0.25, -0.15], [-0.2, 0.0, 0.3]])
    b.set_atoms(1, moved)
    E3m, gradsm = b.adjoint_gradient(spec)
    b.close()
    for m, (rows, NN, screen) in enumerate(mols):
        E, g = _single(rows, NN, screen, n_steps, checkpoint=checkpoint)
        print(f"mol {m}: E batch {E3[m].sum():.12f} single {E:.12f}; max |dE/dR diff| {np.abs(grads[m] - g).max():.2e}")
        assert E3[m].sum() == pytest.approx(E, rel=1e-12)
        assert np.abs(grads[m] - g).max() <= 1e-9 + 1e-9 * np.abs(g).max()
    E, g = _single(mols[1][0], mols[1][1], mols[1][2], n_steps, pos_rows=moved, checkpoint=checkpoint)
    assert E3m[1].sum() == pytest.approx(E, rel=1e-12)
    assert np.abs(gradsm[1] - g).max() <= 1e-9 + 1e-9 * np.abs(g).max()
    # the other molecules did not move
    for m in (0, 2):
        assert np.array_equal(E3m[m], E3[m]) and np.array_equal(gradsm[m], grads[m])


def test_batched_frozen_energies_match_the_gradient_pass():
    mols = _mols()
    b = Batch(mols, mg_levels=3, graph=False)
    b.initialize(bootstrap=0)
    spec = make_spec(20, bootstrap=4, checkpoint=8, poisson_interval=1)
    E_fwd = b.frozen_energies(spec)
    E_adj, _ = b.adjoint_gradient(spec)
    b.close()
    assert np.array_equal(E_fwd, E_adj)
