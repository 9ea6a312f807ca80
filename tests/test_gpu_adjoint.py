"""GPU adjoint forces: hand-written reverse sweep (librq.so) against autograd of
the FP64 torch oracle of the same frozen-label trajectory, against central
finite differences of the GPU solve, checkpoint-interval and tile-shape
invariance, and the torch.autograd.Function wrapper."""
import numpy as np
import pytest
import torch

from oracle import adjoint as OA
from oracle import s1
from realmolecule_amd.adjoint import FrozenSolver, trajectory_energy
from tests.helpers import three_atoms

pytestmark = pytest.mark.gpu


def _oracle_grad(fz, n_steps, **kw):
    """Oracle dE/dR (per bohr, electronic + nuclear) at the solver's FP32 positions
    with the solver's own frozen label map."""
    rows = fz.rows.copy()
    rows[:, :3] = rows[:, :3].astype(np.float32)
    lab = fz.s.label.get()
    E, g = OA.energy_and_gradient(rows, lab, fz.NN, fz.screen, n_steps, **kw)
    g = g - s1.nuclear_forces(rows[:, :3], rows[:, 3], fz.h)
    E += s1.nuclear_energy(rows[:, :3], rows[:, 3], fz.h)
    return E, g


@pytest.mark.parametrize("mode", ["jacobi", "multigrid"])
def test_adjoint_matches_fp64_autograd(mode):
    NN, screen = 16, 16 * 0.3
    rows = three_atoms(NN)
    n_steps = 60
    kw = dict(bootstrap=5, checkpoint=16, poisson_interval=2 if mode == "jacobi" else 1, norm_interval=20)
    extra = {} if mode == "jacobi" else {"poisson": "multigrid", "mg_levels": 3}
    fz = FrozenSolver(rows, NN, screen, n_steps, **kw, **extra)
    pos = rows[:, :3] * fz.h
    E, g = fz.energy_and_gradient(pos)
    okw = dict(bootstrap=5, poisson_interval=kw["poisson_interval"], norm_interval=20)
    if mode == "multigrid":
        okw.update(mode="multigrid", mg_levels=3)
    Eo, go = _oracle_grad(fz, n_steps, **okw)
    err = np.abs(g - go).max()
    print(f"{mode}: E gpu {E:.9f} oracle {Eo:.9f}; max |dE/dR diff| {err:.2e} (|g|max {np.abs(go).max():.3f})")
    assert abs(E - Eo) < 2e-6 * abs(Eo)
    assert err < 1e-4 + 1e-4 * np.abs(go).max()
    fz.close()


def test_adjoint_matches_gpu_central_differences():
    """Bare atoms (rc = 0, so K and the trial are smooth in R); FP32 solves,
    central differences at delta = 5e-3 bohr."""
    NN, screen = 32, 32 * 0.25
    rows = three_atoms(NN)
    rows[:, 4] = 0.0
    fz = FrozenSolver(rows, NN, screen, 100, bootstrap=20, checkpoint=32)
    pos = rows[:, :3] * fz.h
    E, g = fz.energy_and_gradient(pos)
    F_fd = fz.forces_fd(pos, delta=5e-3, atoms=[0, 1])
    diff = np.abs(-g[:2] - F_fd[:2]).max()
    print(f"adjoint vs FD(5e-3 bohr): max diff {diff:.2e} Ha/bohr; F={-g[:2]}")
    assert diff < 1e-3
    fz.close()


def test_rc_clamp_kink_is_excluded_from_the_gradient():
    """Inside rc the potential is clamped (r~ = rc, cloud/solver.py:45-46), so
    voxels there carry no dependence on R: the oracle (torch.clamp) and the
    kernel agree on this convention."""
    NN, screen = 16, 16 * 0.3
    rows = three_atoms(NN)
    rows[1, 4] = 0.9  # many voxels inside the clamp sphere
    fz = FrozenSolver(rows, NN, screen, 20, bootstrap=2, checkpoint=8)
    E, g = fz.energy_and_gradient(rows[:, :3] * fz.h)
    Eo, go = _oracle_grad(fz, 20, bootstrap=2)
    assert np.abs(g - go).max() < 1e-4 + 1e-4 * np.abs(go).max()
    fz.close()


def test_checkpoint_interval_does_not_change_gradient():
    NN, screen = 24, 24 * 0.3
    rows = three_atoms(NN)
    out = []
    for ck in (7, 64, 0):
        fz = FrozenSolver(rows, NN, screen, 50, bootstrap=4, checkpoint=ck)
        out.append(fz.energy_and_gradient(rows[:, :3] * fz.h))
        fz.close()
    assert out[0][0] == pytest.approx(out[1][0], rel=1e-12)
    assert np.allclose(out[0][1], out[1][1], rtol=1e-5, atol=1e-7)
    # one segment recorded by the forward pass (64 >= 50 steps, and the
    # automatic interval) against seven, six of them recomputed
    assert out[2][0] == pytest.approx(out[1][0], rel=1e-12)
    assert np.allclose(out[2][1], out[1][1], rtol=1e-5, atol=1e-7)


@pytest.mark.parametrize("tile", [(128, 2, 4), (32, 8, 16), (256, 1, 3)])
def test_tile_shape_does_not_change_gradient(tile):
    """The forward and reverse sweeps take the fine level's tile (128 x 2 on
    grids of 255 or more nodes per axis): every shape gives the same energy
    and gradient as the default, up to the order of the FP64 block sums."""
    NN, screen = 24, 24 * 0.3
    rows = three_atoms(NN)
    fz = FrozenSolver(rows, NN, screen, 30, bootstrap=4, poisson="multigrid", mg_levels=3)
    pos = rows[:, :3] * fz.h
    E0, g0 = fz.energy_and_gradient(pos)
    fz.s._chk(fz.s._lib.rq_set_tile(fz.s._h, 0, *tile))
    E1, g1 = fz.energy_and_gradient(pos)
    fz.close()
    assert E1 == pytest.approx(E0, rel=1e-12)
    assert np.allclose(g1, g0, rtol=1e-5, atol=1e-7)


def test_autograd_function_backward():
    NN, screen = 16, 16 * 0.3
    rows = three_atoms(NN)
    fz = FrozenSolver(rows, NN, screen, 30, bootstrap=3, checkpoint=8)
    pos = torch.tensor(rows[:, :3] * fz.h, dtype=torch.float64, requires_grad=True)
    E = trajectory_energy(pos, fz)
    E.backward()
    E2, g = fz.energy_and_gradient(rows[:, :3] * fz.h)
    assert float(E.detach()) == pytest.approx(E2, rel=1e-12)
    assert np.allclose(pos.grad.numpy(), g, rtol=1e-6, atol=1e-9)
    fz.close()


def test_relax_h2_on_adjoint_forces():
    """L-BFGS on adjoint gradients lowers the energy of a stretched H2 and
    shrinks its largest force; the first evaluation equals the plain energy."""
    from realmolecule_amd.relax import relax
    NN, screen = 32, 32 * 0.25
    c = NN / 2
    rows = np.array([[c - 5.0, c + 0.3, c, 1.0, 0.0], [c + 5.0, c - 0.2, c + 0.1, 1.0, 0.0]])
    out = relax(rows, NN, screen, n_steps=60, bootstrap=10, max_iter=6, checkpoint=32)
    E = out["energies"]
    assert out["evaluations"] >= 3
    assert min(E) < E[0] - 1e-4
    assert out["grad_max"][-1] < out["grad_max"][0]


def test_saved_start_composes_with_the_cold_map_bitexact():
    """40 cold steps, saved, then 20 steps from the saved state: the energy of
    60 cold steps, bit for bit (same kernels in the same order)."""
    NN, screen = 16, 16 * 0.3
    rows = three_atoms(NN)
    fz = FrozenSolver(rows, NN, screen, 60, bootstrap=5, checkpoint=16)
    pos = rows[:, :3] * fz.h
    E60 = fz.energy(pos)
    E40 = fz.warm_start(pos, 40)
    fz.spec.n_steps = 20
    E_warm = fz.energy(pos)
    print(f"cold 40 {E40:.12f}, cold 60 {E60:.12f}, saved 40 + 20 {E_warm:.12f}")
    assert E_warm == E60
    fz.close()


@pytest.mark.parametrize("mode", ["jacobi", "multigrid"])
def test_saved_start_gradient_matches_fp64_autograd(mode):
    NN, screen = 16, 16 * 0.3
    rows = three_atoms(NN)
    pi = 2 if mode == "jacobi" else 1
    extra = {} if mode == "jacobi" else {"poisson": "multigrid", "mg_levels": 3}
    fz = FrozenSolver(rows, NN, screen, 30, bootstrap=5, checkpoint=8, poisson_interval=pi, **extra)
    pos = rows[:, :3] * fz.h
    fz.warm_start(pos, 60)
    U0, P0 = fz.s.U.get().astype(np.float64), fz.s.P.get().astype(np.float64)
    # a displaced geometry: the saved state stays the one converged at pos
    pos2 = pos.copy()
    pos2[0] += np.array([0.05, -0.03, 0.02])
    E, g = fz.energy_and_gradient(pos2)
    rows2 = rows.copy()
    rows2[:, :3] = (pos2 / fz.h).astype(np.float32)
    lab = fz.s.label.get()
    okw = dict(poisson_interval=pi, norm_interval=20, start=(U0, P0))
    if mode == "multigrid":
        okw.update(mode="multigrid", mg_levels=3)
    Eo, go = OA.energy_and_gradient(rows2, lab, NN, screen, 30, **okw)
    go = go - s1.nuclear_forces(rows2[:, :3], rows2[:, 3], fz.h)
    Eo += s1.nuclear_energy(rows2[:, :3], rows2[:, 3], fz.h)
    err = np.abs(g - go).max()
    print(f"saved start, {mode}: E gpu {E:.9f} oracle {Eo:.9f}; max |dE/dR diff| {err:.2e}")
    assert abs(E - Eo) < 2e-6 * abs(Eo)
    assert err < 1e-4 + 1e-4 * np.abs(go).max()
    fz.close()


@pytest.mark.parametrize("checkpoint", [0, 8])
def test_adjoint_forward_on_the_fine_marches_is_bitexact(checkpoint, monkeypatch):
    """With the fine-level marches in the frozen map's Poisson steps (size
    gate lowered; on by default from 320 nodes per axis), energy and
    gradient equal those of the separate passes bit for bit, whole
    trajectory recorded (checkpoint 0) or recomputed from checkpoints."""
    NN, screen = 96, 96 * 0.12
    rows = three_atoms(NN)
    out = []
    for gate in ("100000", "0"):
        monkeypatch.setenv("RQ_SMOOTH2_MIN_S", gate)
        fz = FrozenSolver(rows, NN, screen, 30, bootstrap=3, checkpoint=checkpoint, poisson_interval=1,
                          poisson="multigrid", mg_levels=4)
        pos = rows[:, :3] * fz.h
        out.append(fz.energy_and_gradient(pos))
        fz.close()
    print(f"separate passes E {out[0][0]:.12f}, marches E {out[1][0]:.12f}")
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1])
