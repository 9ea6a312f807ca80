"""GPU parity of the molecule.js (S2) semantics against the S2 oracle."""
import ctypes as C

import numpy as np
import pytest

from oracle import s1, s2, mg
from tests.helpers import three_atoms, make_solver, cube, ulp_diff

pytestmark = pytest.mark.gpu
F32 = np.float32


def _s2_solver(rows, NN, screen, **kw):
    cfg = dict(init="s2", energy="s2", boundary="levelset", smooth_a=0.3333, smooth_c=9.0)
    cfg.update(kw)
    return make_solver(rows, NN, screen, **cfg)


def test_init_s2_all_nodes():
    rows = three_atoms(24)
    NN, screen = 24, 24 * 0.3
    s = _s2_solver(rows, NN, screen)
    s._chk(s._lib.rq_init_trial(s._h))
    g = s1.Grid(NN, screen, 0.12, F32)
    U, lab, K, _ = s2.init_s2(rows, g)
    S = NN + 1
    assert np.allclose(cube(s.K_buf.get(), S), K, rtol=2e-6)
    assert np.allclose(cube(s.U.get(), S), U, rtol=4e-6, atol=1e-30)
    assert (cube(s.label.get(), S) != lab).sum() <= 1
    assert np.all(cube(s.W.get(), S) == 1.0)
    s.close()


def test_levelset_bitexact():
    rows = three_atoms(24)
    NN, screen = 24, 24 * 0.3
    S = NN + 1
    g = s1.Grid(NN, screen, 0.12, F32)
    U, lab, K, _ = s2.init_s2(rows, g)
    rng = np.random.default_rng(2)
    W = rng.uniform(-0.3, 1.2, (S, S, S)).astype(F32)
    U = (U * rng.uniform(0.5, 1.5, U.shape)).astype(F32)
    for lock in (None, np.array([1, 1, 0], np.int32)):
        s = _s2_solver(rows, NN, screen)
        s.U.set(U)
        s.label.set(lab)
        s.W.set(W)
        if lock is not None:
            s._chk(s._lib.rq_set_lock_groups(s._h, lock.ctypes.data_as(C.POINTER(C.c_int32)), 3))
        s._boundary()
        Lr, Wr = s2.levelset(U, lab, W, g, 0.5, 0.15, lock)
        Lg, Wg = cube(s.label.get(), S), cube(s.W.get(), S)
        assert (Lr != lab).any()
        assert np.array_equal(Lg, Lr)
        assert ulp_diff(Wg, Wr) == 0
        s.close()


def test_energy_s2_matches_oracle():
    rows = three_atoms(24)
    NN, screen = 24, 24 * 0.3
    S = NN + 1
    g = s1.Grid(NN, screen, 0.12, F32)
    U, lab, K, _ = s2.init_s2(rows, g)
    P = np.random.default_rng(3).uniform(0, 1, (S, S, S)).astype(F32)
    s = _s2_solver(rows, NN, screen)
    s.U.set(U)
    s.K_buf.set(K)
    s.P.set(P)
    s.label.set(lab)
    s._compute_energy()
    tot, per, dip = s2.energy_s2(U, K, P, lab, rows, g)
    assert np.allclose([s.E_T, s.E_eK, s.E_ee], tot, rtol=1e-12)
    assert np.allclose(s.E_domains, per, rtol=1e-12, atol=1e-15)
    assert np.allclose(s.dipole, dip, rtol=1e-12)
    s.close()


def test_gradp_forces_match_oracle():
    rows = three_atoms(32)
    NN, screen = 32, 32 * 0.25
    S = NN + 1
    g = s1.Grid(NN, screen, 0.12, F32)
    P = np.random.default_rng(4).uniform(0, 1, (S, S, S)).astype(F32)
    s = _s2_solver(rows, NN, screen)
    s.P.set(P)
    got = s.forces_gradp(6)
    ref = s2.gradp_forces(P, rows, g, R=6)
    assert np.allclose(got, ref, rtol=1e-10, atol=1e-12)
    s.close()


def test_itp_rc_dirichlet_bitexact():
    rows = three_atoms(24)
    rows[1, 4] = 0.9
    NN, screen = 24, 24 * 0.3
    S = NN + 1
    g = s1.Grid(NN, screen, s1.dv_table(3), F32)
    U, lab, K, _ = s2.init_s2(rows, g)
    P = np.random.default_rng(5).uniform(0, 1, (S, S, S)).astype(F32)
    s = _s2_solver(rows, NN, screen, rc_dirichlet=True)
    s.U.set(U)
    s.K_buf.set(K)
    s.P.set(P)
    s.label.set(lab)  # rc flags are derived on upload
    s._update()
    mask = s2.rc_mask(lab, rows, g)
    assert mask.sum() > 0
    ref = s1.itp_update(U, K, P, lab, g, rc_mask=mask)
    assert ulp_diff(cube(s.U.get(), S), ref) == 0
    s.close()


def test_twogrid_poisson_step():
    """Two damped sweeps + the sign-inverted two-grid correction
    (molecule.js:2869-2907), one Poisson step at a correction step."""
    rows = three_atoms(32)
    NN, screen = 32, 32 * 0.25
    S = NN + 1
    g = s1.Grid(NN, screen, 0.12, F32)
    rng = np.random.default_rng(6)
    U = np.zeros((S, S, S), F32)
    U[1:-1, 1:-1, 1:-1] = rng.uniform(0, 1, (S - 2,) * 3)
    P = np.zeros((S, S, S), F32)
    P[1:-1, 1:-1, 1:-1] = rng.uniform(0, 1, (S - 2,) * 3)
    s = _s2_solver(rows, NN, screen, poisson="twogrid", twogrid_interval=1, boundary="off")
    s.U.set(U)
    s.P.set(P)
    s.step(0, 1, poisson_interval=1, norm_interval=10 ** 6)
    rhs = s1.rho_rhs(U, g.two_pi, g.h2, F32)
    ref = s1.smooth(P, rhs, 0.3333, 9.0, F32)
    ref = s1.smooth(ref, rhs, 0.3333, 9.0, F32)
    ref = mg.twogrid_s2(ref, rhs, F32)
    got = cube(s.P.get(), S)
    assert ulp_diff(got, ref) == 0
    s.close()


def test_pother_direct_matches_oracle():
    """Direct per-domain potential (molecule.js direct Pother): two Poisson
    steps (one V-cycle per domain each, persistent P_m) and the label gather,
    against the FP32 oracle; then one ITP update with that potential."""
    rows = three_atoms(32)
    NN, screen = 32, 32 * 0.25
    S = NN + 1
    g = s1.Grid(NN, screen, 0.12, F32)
    U, lab, K, _ = s1.init_trial(rows, g)
    U, _ = s1.normalize(U, lab, rows[:, 3], g)
    s = make_solver(rows, NN, screen, poisson="multigrid", mg_levels=3, pother="direct")
    s.U.set(U)
    s.label.set(lab)
    s._poisson_step()
    s._poisson_step()
    got = cube(s.P.get(), S)
    Pdom = [np.zeros((S, S, S), F32) for _ in range(len(rows))]
    P, Pdom = s2.pother_direct(U, lab, Pdom, rows[:, 3], g, 3, 2, 2, 1 / 7, 7.0)
    P, Pdom = s2.pother_direct(U, lab, Pdom, rows[:, 3], g, 3, 2, 2, 1 / 7, 7.0)
    err = np.abs(got - P).max() / np.abs(P).max()
    print(f"direct Pother max rel diff {err:.2e}, ulp {ulp_diff(got, P)}")
    assert err < 1e-6
    s.K_buf.set(K)
    s._update()
    ref = s1.itp_update(U, K, got, lab, g)
    assert ulp_diff(cube(s.U.get(), S), ref) == 0
    s.close()


def test_pother_direct_rejected_without_multigrid():
    rows = three_atoms(16)
    with pytest.raises(Exception):
        make_solver(rows, 16, 16 * 0.3, poisson="jacobi", pother="direct")


@pytest.mark.parametrize("sk", [16, 1, 31])
def test_display_slice_bitexact(sk):
    rows = three_atoms(32)
    NN, screen = 32, 32 * 0.25
    S = NN + 1
    rng = np.random.default_rng(sk)
    U = rng.uniform(0.0, 0.01, (S, S, S)).astype(F32)
    U[rng.random((S, S, S)) < 0.3] = 0.0
    lab = rng.integers(0, 3, (S, S, S)).astype(np.int32)
    K = rng.uniform(0, 5, (S, S, S)).astype(F32)
    s = make_solver(rows, NN, screen)
    s.U.set(U)
    s.label.set(lab)
    s.K_buf.set(K)
    got = s.display_slice(sk)["raw"]
    ref = s2.display_slice(U, lab, K, rows, sk, NN)
    assert np.array_equal(got, ref)
    s.close()


def test_pother_frame_sic_matches_oracle():
    """molecule.js frame potential (PotherBuf, :3009-3076): Pother = P after
    initialisation; after a SIC frame it is P minus each domain's
    self-potential on its own nodes, bit-exact against the oracle on the
    device's own fields; the next step's ITP and the energy read it."""
    rows = three_atoms(32)
    NN, screen = 32, 32 * 0.25
    S = NN + 1
    g = s1.Grid(NN, screen, 0.12, F32)
    s = make_solver(rows, NN, screen, poisson="twogrid", twogrid_interval=2, smooth_a=0.3333, smooth_c=9.0,
                    pother="frame", sic_interval=2, bootstrap_poisson=4)
    s.initialize()
    assert ulp_diff(cube(s.Pother.get(), S), cube(s.P.get(), S)) == 0
    s.step(0, 6, poisson_interval=1)
    U, P, K = cube(s.U.get(), S), cube(s.P.get(), S), cube(s.K_buf.get(), S)
    lab = cube(s.label.get(), S)
    s.pother_frame(1)  # not a SIC frame: a copy of P
    assert ulp_diff(cube(s.Pother.get(), S), P) == 0
    s.pother_frame(2)
    got = cube(s.Pother.get(), S)
    ref = s2.pother_frame(U, lab, P, rows[:, 3], g, 2, 2, 0.3333, 9.0)
    assert ulp_diff(got, ref) == 0
    assert np.abs(got - P).max() > 0
    # the ITP update reads Pother (one step without a Poisson step or normalisation)
    s.step(6, 1, poisson_interval=10 ** 6, norm_interval=10 ** 6, boundary_interval=10 ** 6)
    assert ulp_diff(cube(s.U.get(), S), s1.itp_update(U, K, ref, lab, g)) == 0
    s._compute_energy()
    tot, _ = s1.energy(cube(s.U.get(), S), K, ref, lab, g, 3)
    assert np.allclose([s.E_T, s.E_eK, s.E_ee], tot, rtol=1e-12)
    s.close()


def test_pother_frame_sic_needs_a_coarse_level():
    rows = three_atoms(16)
    with pytest.raises(Exception):
        make_solver(rows, 16, 16 * 0.3, poisson="jacobi", pother="frame", sic_interval=1)
    with pytest.raises(Exception):
        make_solver(rows, 16, 16 * 0.3, poisson="twogrid", sic_interval=1)  # SIC without the frame mode
