"""GPU parity of the cloud/solver.py (S1) passes against the CPU oracle.

Each test writes the same inputs to the device and to the oracle, runs one
pass through the C ABI and compares.  Kernel-level passes built only from
+ - * / and sqrt are bit-exact against the oracle's FP32-faithful mode; passes
with transcendental functions (expf in the trial init) or FP64 reductions are
checked to a stated tolerance, and the FP64-oracle error is printed.
"""
import numpy as np
import pytest

from oracle import s1, mg
from tests.helpers import small_h2o, three_atoms, make_solver, cube, ulp_diff, rel_l2

pytestmark = pytest.mark.gpu

F32 = np.float32


def _state(rows, NN, screen, seed=0, dtype=F32):
    """Oracle inputs after init + one normalise, plus random P."""
    g = s1.Grid(NN, screen, s1.dv_table(int((rows[:, 3] > 0).sum())), dtype)
    U, lab, K, _ = s1.init_trial(rows, g)
    rng = np.random.default_rng(seed)
    S = NN + 1
    P = np.zeros((S, S, S), dtype)
    P[1:-1, 1:-1, 1:-1] = rng.uniform(0.0, 2.0, (S - 2,) * 3).astype(dtype)
    return g, U, lab, K, P


def test_init_trial_matches_oracle():
    rows, els, screen = small_h2o(40, 0.2)
    NN = 40
    s = make_solver(rows, NN, screen)
    s._chk(s._lib.rq_init_trial(s._h))
    g = s1.Grid(NN, screen, 0.12, F32)
    U, lab, K, _ = s1.init_trial(rows, g)
    S = NN + 1
    Ug, Kg, Lg = cube(s.U.get(), S), cube(s.K_buf.get(), S), cube(s.label.get(), S)
    print(f"init: K max rel {np.abs(Kg - K).max() / np.abs(K).max():.2e}, U max ulp {ulp_diff(Ug, U)}")
    assert np.allclose(Kg, K, rtol=2e-6, atol=0)
    assert np.allclose(Ug, U, rtol=4e-6, atol=1e-30)
    # labels identical except where two trial values tie to within float rounding
    diff = Lg != lab
    assert diff.sum() <= max(1, int(1e-5 * diff.size)), diff.sum()
    s.close()


def test_itp_update_bitexact():
    rows = three_atoms(24)
    NN, screen = 24, 24 * 0.3
    g, U, lab, K, P = _state(rows, NN, screen)
    s = make_solver(rows, NN, screen)
    s.U.set(U)
    s.K_buf.set(K)
    s.P.set(P)
    s.label.set(lab)
    s._update()
    got = cube(s.U.get(), NN + 1)
    ref = s1.itp_update(U, K, P, lab, g)
    assert ulp_diff(got, ref) == 0
    s.close()


def test_jacobi_bitexact():
    rows = three_atoms(24)
    NN, screen = 24, 24 * 0.3
    g, U, lab, K, P = _state(rows, NN, screen, seed=1)
    s = make_solver(rows, NN, screen)
    s.U.set(U)
    s.P.set(P)
    s.label.set(lab)
    s._poisson_step()
    got = cube(s.P.get(), NN + 1)
    ref = s1.jacobi(P, U, g, 2)
    assert ulp_diff(got, ref) == 0
    s.close()


def test_normalize_matches_oracle():
    rows = three_atoms(24)
    NN, screen = 24, 24 * 0.3
    g, U, lab, K, P = _state(rows, NN, screen)
    s = make_solver(rows, NN, screen)
    s.U.set(U)
    s.label.set(lab)
    norms_gpu = s.domain_norms()
    norms_ref = s1.domain_norms(U, lab, 3, g)
    assert np.allclose(norms_gpu, norms_ref, rtol=1e-12, atol=0)
    s._normalize()
    got = cube(s.U.get(), NN + 1)
    ref, _ = s1.normalize(U, lab, rows[:, 3], g)
    assert ulp_diff(got, ref) <= 1
    # invariant: sum_{label=d} U^2 h^3 = Z_d after normalise
    after = s.domain_norms()
    assert np.allclose(after, rows[:, 3], rtol=1e-6)
    s.close()


def test_boundary_score_exact():
    rows = three_atoms(24)
    NN, screen = 24, 24 * 0.3
    g, U, lab, K, P = _state(rows, NN, screen)
    rng = np.random.default_rng(5)
    # perturb labels near walls so the rule has flips to make
    lab2 = lab.copy()
    inner = lab2[1:-1, 1:-1, 1:-1]
    m = rng.random(inner.shape) < 0.1
    inner[m] = rng.integers(0, 3, m.sum())
    s = make_solver(rows, NN, screen)
    s.U.set(U)
    s.label.set(lab2)
    s._boundary()
    got = cube(s.label.get(), NN + 1)
    ref = s1.boundary_score(U, lab2, rows, g)
    assert (got != lab2).sum() > 0
    assert np.array_equal(got, ref)
    s.close()


@pytest.mark.parametrize("n_atoms,NN,h", [(60, 40, 0.35), (150, 48, 0.6), (300, 48, 0.3), (1500, 24, 0.1)])
def test_boundary_score_many_atoms_cell_search(n_atoms, NN, h):
    """Many atoms (mixed Z, one with Z = 0): the per-node best atom (from
    the block-candidate map, or the kernel's own cell-list search on blocks
    with more than 512 candidates, as 1500 atoms packed in a 2.4-bohr box
    give) must be the same atom as the oracle's scan over all atoms,
    including boundary voxels far from every nucleus."""
    rng = np.random.default_rng(11 + n_atoms)
    rows = np.zeros((n_atoms, 5))
    rows[:, :3] = rng.uniform(2.0, NN - 2.0, (n_atoms, 3))
    rows[:, 3] = rng.choice([1.0, 2.0, 3.0, 4.0], n_atoms)
    rows[3, 3] = 0.0
    screen = NN * h
    g, U, lab, K, P = _state(rows, NN, screen)
    lab2 = lab.copy()
    inner = lab2[1:-1, 1:-1, 1:-1]
    m = rng.random(inner.shape) < 0.2
    inner[m] = rng.integers(0, n_atoms, m.sum())
    s = make_solver(rows, NN, screen)
    s.U.set(U)
    s.label.set(lab2)
    s._boundary()
    got = cube(s.label.get(), NN + 1)
    ref = s1.boundary_score(U, lab2, rows, g)
    assert (got != lab2).sum() > 0
    assert np.array_equal(got, ref)
    s.close()


@pytest.mark.parametrize("n_atoms,NN,h", [(3, 24, 0.3), (300, 48, 0.3)])
def test_boundary_score_best_atom_cache(n_atoms, NN, h):
    """The per-node best atom kept between applications of the score rule:
    two applications in a row (the second reads the atoms found by the
    first), then an atom moved (rq_set_atoms must drop every kept entry) and
    a third application, each against the oracle's scan over all atoms on
    the same labels; 300 atoms take the cell-search kernel."""
    rng = np.random.default_rng(23 + n_atoms)
    if n_atoms == 3:
        rows = three_atoms(NN)
    else:
        rows = np.zeros((n_atoms, 5))
        rows[:, :3] = rng.uniform(2.0, NN - 2.0, (n_atoms, 3))
        rows[:, 3] = rng.choice([1.0, 2.0, 3.0], n_atoms)
    screen = NN * h
    g, U, lab, K, P = _state(rows, NN, screen)
    lab1 = lab.copy()
    inner = lab1[1:-1, 1:-1, 1:-1]
    m = rng.random(inner.shape) < 0.2
    inner[m] = rng.integers(0, n_atoms, m.sum())
    s = make_solver(rows, NN, screen)
    s.U.set(U)
    s.label.set(lab1)
    s._boundary()
    got1 = cube(s.label.get(), NN + 1)
    ref1 = s1.boundary_score(U, lab1, rows, g)
    assert np.array_equal(got1, ref1)
    s._boundary()
    got2 = cube(s.label.get(), NN + 1)
    assert np.array_equal(got2, s1.boundary_score(U, got1, rows, g))
    # move atom 1 across the grid: its best-atom regions change
    rows2 = rows.copy()
    rows2[1, :3] = rows[0, :3] + np.array([2.0, -3.0, 1.5])
    s.atom_pos[1][:] = rows2[1, :3]
    s._chk(s._lib.rq_set_atoms(s._h, s._atoms_c(), s.nAtoms))
    lab3 = lab1.copy()
    s.label.set(lab3)
    s._boundary()
    got3 = cube(s.label.get(), NN + 1)
    ref3 = s1.boundary_score(U, lab3, rows2, g)
    assert (ref3 != s1.boundary_score(U, lab3, rows, g)).sum() > 0
    assert np.array_equal(got3, ref3)
    s.close()


def test_energy_matches_oracle():
    rows = three_atoms(24)
    NN, screen = 24, 24 * 0.3
    g, U, lab, K, P = _state(rows, NN, screen)
    s = make_solver(rows, NN, screen)
    s.U.set(U)
    s.K_buf.set(K)
    s.P.set(P)
    s.label.set(lab)
    s._compute_energy()
    tot, per = s1.energy(U, K, P, lab, g, 3)
    got = np.array([s.E_T, s.E_eK, s.E_ee])
    assert np.allclose(got, tot, rtol=1e-12)
    assert np.allclose(s.E_domains, per, rtol=1e-12, atol=1e-14)
    # the solver keeps atom_pos in FP32, as cloud/solver.py:335 does
    ekk = s1.nuclear_energy(rows[:, :3].astype(np.float32), rows[:, 3], g.h64)
    assert abs(s.E_KK - ekk) < 1e-12 * abs(ekk)
    s.close()


def test_forces_hf_match_oracle():
    rows = three_atoms(24)
    NN, screen = 24, 24 * 0.3
    g, U, lab, K, P = _state(rows, NN, screen)
    s = make_solver(rows, NN, screen)
    s.U.set(U)
    out = np.zeros(9)
    import ctypes as C
    buf = (C.c_double * 9)()
    s._chk(s._lib.rq_forces_hf(s._h, 0, 3, buf))
    got = np.frombuffer(buf, np.float64).reshape(3, 3)
    ref = s1.hf_forces(U, rows, g)
    assert np.allclose(got, ref, rtol=1e-9, atol=1e-12)
    s.close()


def test_multigrid_vcycle_matches_oracle():
    rows = three_atoms(32)
    NN, screen = 32, 32 * 0.25
    g, U, lab, K, P = _state(rows, NN, screen, seed=3)
    s = make_solver(rows, NN, screen, poisson="multigrid", mg_levels=3)
    s.U.set(U)
    s.P.set(P)
    s.label.set(lab)
    s._poisson_step()
    got = cube(s.P.get(), NN + 1)
    a, c = np.float32(1.0 / 7.0), np.float32(7.0)
    rhs = s1.rho_rhs(U, g.two_pi, g.h2, F32)
    ref = mg.vcycle(P, rhs, 3, 2, 2, float(a), float(c), F32)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"vcycle max rel diff {err:.2e}, ulp {ulp_diff(got, ref)}")
    assert err < 1e-6


def test_vcycle_converges_to_direct_solve():
    rows = three_atoms(32)
    NN, screen = 32, 32 * 0.25
    g, U, lab, K, P = _state(rows, NN, screen, seed=3)
    s = make_solver(rows, NN, screen, poisson="multigrid", mg_levels=4)
    s.U.set(U)
    s.P.set(np.zeros_like(P))
    for _ in range(12):
        s._poisson_step()
    got = cube(s.P.get(), NN + 1)[1:-1, 1:-1, 1:-1]
    rhs = s1.rho_rhs(U.astype(np.float64), 2 * np.pi, (screen / NN) ** 2, np.float64)
    exact = mg.spsolve_poisson(rhs)
    err = rel_l2(got, exact)
    print(f"12 V-cycles vs direct solve: rel L2 {err:.2e}")
    assert err < 2e-6


def test_trajectory_s1_matches_fp32_oracle():
    """200 steps of the cloud/solver.py schedule from initialize(); density L2
    and energies against the FP32-faithful oracle (tolerances of the parity
    bar), FP64-oracle error reported."""
    rows, els, screen = small_h2o(40, 0.2)
    NN = 40
    s = make_solver(rows, NN, screen)
    s.initialize()
    s.step(0, 200)
    s._compute_energy()
    st32 = s1.State(rows, NN, screen, F32)
    st32.initialize()
    st64 = s1.State(rows, NN, screen, np.float64)
    st64.initialize()
    for step in range(1, 201):
        st32.step(step)
        st64.step(step)
    S = NN + 1
    Ug = cube(s.U.get(), S)
    rho_err = rel_l2(Ug * Ug, st32.U.astype(np.float64) ** 2)
    tot32, per32, ekk = st32.energy()
    tot64, per64, _ = st64.energy()
    E_gpu = s.E
    E32 = tot32.sum() + ekk
    E64 = tot64.sum() + ekk
    print(f"rho L2 vs FP32 oracle {rho_err:.2e}; E gpu {E_gpu:.9f} fp32 {E32:.9f} fp64 {E64:.9f}; "
          f"rel vs fp32 {abs(E_gpu - E32) / abs(E32):.2e} vs fp64 {abs(E_gpu - E64) / abs(E64):.2e}")
    assert rho_err < 1e-5
    assert abs(E_gpu - E32) <= 1e-6 * abs(E32)
    assert np.allclose(s.E_domains, per32, rtol=1e-6, atol=1e-7)
    assert np.array_equal(cube(s.label.get(), S), st32.label)
    s.close()


def test_recompute_k_bitexact_after_move():
    """rq_set_atoms + rq_recompute_k (cloud/solver.py:538-548, 149-179)
    against the FP32 oracle's K refresh at the moved positions: bit-exact."""
    rows = three_atoms(32)
    NN, screen = 32, 32 * 0.25
    s = make_solver(rows, NN, screen)
    s.initialize()
    moved = rows.copy()
    moved[:, :3] += np.array([[0.37, -0.21, 0.05], [-0.13, 0.29, -0.41], [0.08, 0.11, -0.33]])
    from realmolecule_amd.solver import atoms_array
    s._chk(s._lib.rq_set_atoms(s._h, atoms_array(moved), len(moved)))
    s._chk(s._lib.rq_recompute_k(s._h))
    got = cube(s.K_buf.get(), NN + 1)
    ref = s1.recompute_k(moved, s1.Grid(NN, screen, 0.12, F32))
    print(f"recompute_k: max ulp {ulp_diff(got, ref)}")
    assert ulp_diff(got, ref) == 0
    s.close()


def test_run_with_dynamics_matches_oracle_schedule():
    """RealQMSolver.run() with dynamics on (cloud/solver.py:550-634): 100
    steps, forces + damped nuclear move + K refresh every 50 steps, a
    report every 50 steps.  The snapshots carry the reference's keys; their
    energies and nucleus positions follow the FP32 oracle schedule
    (oracle/s1.py State, hf_forces + nuclear_forces, move_nuclei,
    recompute_k)."""
    rows, els, screen = small_h2o(32, 0.25)
    rows0 = rows.copy()
    NN = 32
    s = make_solver(rows, NN, screen, dynamics=True, forceScale=500.0, force_interval=50, bootstrap_poisson=20)
    snaps = s.run(100, report_interval=50)
    keys = {"step", "E", "E_T", "E_eK", "E_ee", "E_KK", "E_bind", "ms_per_step", "elapsed_sec", "nucPos"}
    assert [sn["step"] for sn in snaps] == [50, 100]
    assert all(set(sn) == keys for sn in snaps)

    st = s1.State(rows0.copy(), NN, screen, F32)
    st.initialize(bootstrap=20)
    pos = rows0[:, :3].astype(np.float64).copy()
    vel = np.zeros_like(pos)
    Z = rows[:, 3]
    h = screen / NN
    n_el = int((Z > 0).sum())
    for step in range(1, 101):
        st.step(step)
        if step % 50 == 0:
            F = s1.hf_forces(st.U, st.atoms, st.g) + s1.nuclear_forces(pos, Z, h)
            pos, vel = s1.move_nuclei(pos, vel, F, Z, h, NN, 0.8, 500.0, 0.98, 0.3 if n_el <= 5 else 0.1)
            st.atoms[:, :3] = pos
            st.K = s1.recompute_k(st.atoms, st.g)
            tot, _, ekk = st.energy()
            sn = snaps[step // 50 - 1]
            E_ref = tot.sum() + ekk
            dpos = np.abs(np.array(sn["nucPos"]) - pos).max()
            print(f"step {step}: E gpu {sn['E']:.9f} oracle {E_ref:.9f}; max |dpos| {dpos:.2e} grid units; "
                  f"moved {np.abs(pos - rows0[:, :3]).max():.3f}")
            assert dpos < 1e-4
            assert abs(sn["E"] - E_ref) <= 1e-6 * abs(E_ref)
            assert abs(sn["E_KK"] - ekk) <= 1e-6 * abs(ekk)
    assert np.abs(pos - rows0[:, :3]).max() > 1e-3  # the nuclei did move
    s.close()
