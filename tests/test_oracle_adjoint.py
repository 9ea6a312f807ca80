"""Pins of the differentiable trajectory oracle (oracle/adjoint.py), CPU only.

trajectory_energy is a torch FP64 rewrite of the frozen-label S1 trajectory
(trial init at the frozen labels, normalise, bootstrap Poisson steps, then
the cloud/solver.py loop body without the boundary rule).  It is the ground
truth of the HIP adjoint tests, so it is checked here against:
  * the NumPy oracle (oracle/s1.py State and passes, oracle/mg.py vcycle),
    which are pinned in tests/test_oracle_s1.py and tests/test_oracle_mg.py,
    on the same trajectory, in Jacobi and multigrid mode, to 1e-12;
  * its own autograd gradient against central differences, to 1e-7 Ha/bohr.
It also pins oracle.s1.hf_forces by autograd: the Hellmann-Feynman force is
-dV_eK/dR at fixed U.  The saved-start map (``start``) is pinned by
composition: n1 cold steps, then n2 steps from the state they reach, give
the energy of n1 + n2 cold steps, bit for bit; its gradient holds the K(R)
part only (central differences of itself).
"""
import numpy as np
import pytest
import torch

from oracle import adjoint as A
from oracle import mg, s1
from tests.helpers import three_atoms


def _numpy_trajectory(rows, NN, screen, n_steps, bootstrap, mode, levels=3, nu=2, a=1.0 / 7.0, c=7.0,
                      poisson_interval=2, norm_interval=20):
    """The same map with the NumPy oracle's passes (FP64)."""
    st = s1.State(rows, NN, screen, np.float64, dv=0.12)
    g = st.g
    U, label, K, _ = s1.init_trial(rows, g)
    U, _ = s1.normalize(U, label, st.Z, g)
    P = np.zeros_like(U)

    def poisson(P, U):
        if mode == "jacobi":
            return s1.jacobi(P, U, g, 2)
        rhs = s1.rho_rhs(U, g.two_pi, g.h2, np.float64)
        return mg.vcycle(P, rhs, levels, nu, nu, a, c, np.float64)

    for _ in range(bootstrap):
        P = poisson(P, U)
    for step in range(1, n_steps + 1):
        if step % poisson_interval == 0:
            P = poisson(P, U)
        U = s1.itp_update(U, K, P, label, g)
        if step % norm_interval == 0:
            U, _ = s1.normalize(U, label, st.Z, g)
    tot, _ = s1.energy(U, K, P, label, g)
    return tot, label


@pytest.mark.parametrize("mode", ["jacobi", "multigrid"])
def test_trajectory_energy_matches_numpy_oracle(mode):
    NN, screen = 16, 16 * 0.3
    rows = three_atoms(NN)
    tot, label = _numpy_trajectory(rows, NN, screen, 40, 6, mode)
    R = torch.tensor(rows[:, :3], dtype=torch.float64)
    E, T, VeK, Vee = A.trajectory_energy(R, rows, label, NN, screen, 40, bootstrap=6, mode=mode, mg_levels=3,
                                         dv=0.12)
    got = np.array([float(T), float(VeK), float(Vee)])
    assert np.allclose(got, tot, rtol=1e-12, atol=0), (got, tot)
    assert abs(float(E) - tot.sum()) <= 1e-12 * abs(tot.sum())


def test_trajectory_energy_labels_are_frozen_at_the_given_map():
    """A different (frozen) label map changes the energy: the map is an input,
    not recomputed from R."""
    NN, screen = 16, 16 * 0.3
    rows = three_atoms(NN)
    lab = A.frozen_labels(rows, NN, screen)
    R = torch.tensor(rows[:, :3], dtype=torch.float64)
    E0 = float(A.trajectory_energy(R, rows, lab, NN, screen, 10, bootstrap=2)[0])
    lab2 = lab.copy()
    S = NN + 1
    L = lab2.reshape(S, S, S)
    L[1:-1, 1:-1, 1:NN // 2] = 0
    E1 = float(A.trajectory_energy(R, rows, lab2, NN, screen, 10, bootstrap=2)[0])
    assert abs(E1 - E0) > 1e-3


@pytest.mark.parametrize("mode", ["jacobi", "multigrid"])
def test_autograd_matches_central_differences(mode):
    NN, screen = 12, 12 * 0.35
    c = NN / 2
    rows = np.array([[c - 2.3, c + 0.4, c + 0.2, 1.0, 0.0],
                     [c + 0.3, c - 0.2, c - 0.5, 2.0, 0.0],
                     [c + 2.6, c + 1.1, c + 0.7, 1.0, 0.0]])
    lab = A.frozen_labels(rows, NN, screen)
    kw = dict(bootstrap=4, mode=mode, mg_levels=2, dv=0.12)
    E, g = A.energy_and_gradient(rows, lab, NN, screen, 24, **kw)
    h = screen / NN
    delta = 1e-4  # bohr
    fd = np.zeros_like(g)
    for a_ in range(3):
        for d in range(3):
            Rp = torch.tensor(rows[:, :3], dtype=torch.float64)
            Rm = Rp.clone()
            Rp[a_, d] += delta / h
            Rm[a_, d] -= delta / h
            Ep = float(A.trajectory_energy(Rp, rows, lab, NN, screen, 24, **kw)[0])
            Em = float(A.trajectory_energy(Rm, rows, lab, NN, screen, 24, **kw)[0])
            fd[a_, d] = (Ep - Em) / (2 * delta)
    assert np.max(np.abs(fd - g)) < 1e-7, np.max(np.abs(fd - g))
    # a translation of all atoms changes E only through the grid: the total
    # force is small next to the per-atom forces, not zero
    assert np.abs(g).max() > 1e-2


def test_hf_forces_are_minus_gradient_of_electron_nucleus_energy():
    """oracle.s1.hf_forces (the cloud/solver.py:282-315 force) is -dV_eK/dR at
    fixed U for rc = 0 atoms, whose K uses the same kernel 1/sqrt(r^2 + h^2):
    autograd of the torch V_eK built as trajectory_energy builds K."""
    NN, screen = 12, 12 * 0.35
    c = NN / 2
    rows = np.array([[c - 2.1, c + 0.4, c + 0.2, 1.0, 0.0], [c + 2.2, c - 0.3, c - 0.4, 2.0, 0.0]])
    st = s1.State(rows, NN, screen, np.float64, dv=0.12)
    U, label, K, _ = s1.init_trial(rows, st.g)
    U, _ = s1.normalize(U, label, st.Z, st.g)
    R = torch.tensor(rows[:, :3], dtype=torch.float64, requires_grad=True)
    h = screen / NN
    x = torch.arange(1, NN, dtype=torch.float64) * h
    Kin = 0
    for n in range(2):
        dx = (x - R[n, 0] * h)[:, None, None]
        dy = (x - R[n, 1] * h)[None, :, None]
        dz = (x - R[n, 2] * h)[None, None, :]
        Kin = Kin + rows[n, 3] / torch.sqrt(dx * dx + dy * dy + dz * dz + h * h)
    assert np.allclose(Kin.detach().numpy(), K[1:-1, 1:-1, 1:-1], rtol=1e-14)
    u = torch.tensor(U[1:-1, 1:-1, 1:-1])
    VeK = -(Kin * u * u).sum() * h ** 3
    VeK.backward()
    dVdR = R.grad.numpy() / h
    F = s1.hf_forces(U, rows, st.g)
    assert np.allclose(F, -dVdR, rtol=1e-10, atol=1e-13), (F, -dVdR)
    assert np.abs(F).max() > 1e-3


@pytest.mark.parametrize("mode", ["jacobi", "multigrid"])
def test_saved_start_composes_with_the_cold_map(mode):
    NN, screen = 12, 12 * 0.35
    c = NN / 2
    rows = np.array([[c - 2.3, c + 0.4, c + 0.2, 1.0, 0.0], [c + 0.3, c - 0.2, c - 0.5, 2.0, 0.0],
                     [c + 2.6, c + 1.1, c + 0.7, 1.0, 0.0]])
    lab = A.frozen_labels(rows, NN, screen)
    kw = dict(mode=mode, mg_levels=2, dv=0.12)
    R = torch.tensor(rows[:, :3], dtype=torch.float64)
    st = {}
    A.trajectory_energy(R, rows, lab, NN, screen, 20, bootstrap=3, state_out=st, **kw)
    E_cold = float(A.trajectory_energy(R, rows, lab, NN, screen, 34, bootstrap=3, **kw)[0])
    E_warm = float(A.trajectory_energy(R, rows, lab, NN, screen, 14, start=(st["U"], st["P"]), **kw)[0])
    assert E_warm == E_cold
    # the gradient of the saved-start map against central differences of it
    E, g = A.energy_and_gradient(rows, lab, NN, screen, 14, start=(st["U"], st["P"]), **kw)
    h = screen / NN
    delta = 1e-4
    for a_, d in ((0, 0), (1, 2), (2, 1)):
        Rp = R.clone()
        Rm = R.clone()
        Rp[a_, d] += delta / h
        Rm[a_, d] -= delta / h
        Ep = float(A.trajectory_energy(Rp, rows, lab, NN, screen, 14, start=(st["U"], st["P"]), **kw)[0])
        Em = float(A.trajectory_energy(Rm, rows, lab, NN, screen, 14, start=(st["U"], st["P"]), **kw)[0])
        assert abs((Ep - Em) / (2 * delta) - g[a_, d]) < 1e-7
    # and it differs from the cold map's gradient (the start's R dependence is gone)
    _, gc = A.energy_and_gradient(rows, lab, NN, screen, 34, bootstrap=3, **kw)
    assert np.abs(gc - g).max() > 1e-4
