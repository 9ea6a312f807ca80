"""Pin the S1 oracle (cloud/solver.py restatement) against independent checks.

Each test checks an oracle function against something other than its own
formula: scalar brute force written as explicit loops, an independently
assembled sparse operator, a direct solve, closed-form eigenvalues, finite
differences of an energy, or a conservation law.  Sizes are CPU-small.
"""
import math

import numpy as np
import pytest

from oracle import s1
from oracle.mg import spsolve_poisson

F64 = np.float64


def atoms3(NN):
    c = NN / 2
    return np.array([[c - 2.3, c + 0.7, c + 0.2, 1.0, 0.0],
                     [c + 0.4, c - 0.3, c - 0.6, 2.0, 0.4],
                     [c + 2.1, c + 1.9, c + 0.9, 1.0, 0.0]])


def test_dv_table():
    assert [s1.dv_table(n) for n in (1, 30, 31, 100, 101, 200, 201, 500, 501)] == \
        [0.12, 0.12, 0.03, 0.03, 0.02, 0.02, 0.01, 0.01, 0.005]


def test_init_trial_scalar_bruteforce():
    NN, screen = 12, 12 * 0.3
    rows = atoms3(NN)
    g = s1.Grid(NN, screen, 0.12, F64)
    U, lab, K, _ = s1.init_trial(rows, g)
    h = screen / NN
    rng = np.random.default_rng(0)
    for _ in range(60):
        i, j, k = rng.integers(1, NN, 3)
        Kv, best, bestn = 0.0, 0.0, 0
        for n, (ai, aj, ak, Z, rc) in enumerate(rows):
            r2 = ((i - ai) * h) ** 2 + ((j - aj) * h) ** 2 + ((k - ak) * h) ** 2
            r = max(math.sqrt(r2 + 0.04 * h * h), rc) if rc > 0 else math.sqrt(r2 + h * h)
            Kv += Z / r
            u = Z * Z * 0.5641895835 * math.exp(-Z * r)
            if u > best:
                best, bestn = u, n
        assert abs(K[i, j, k] - Kv) < 1e-12 * Kv
        assert abs(U[i, j, k] - best) < 1e-12 * max(best, 1e-300)
        assert lab[i, j, k] == bestn
    # faces are zero
    assert not U[0].any() and not K[:, -1].any() and not lab[:, :, 0].any()


def test_init_trial_tie_goes_to_first_atom():
    NN = 10
    rows = np.array([[3.0, 5.0, 5.0, 1.0, 0.0], [7.0, 5.0, 5.0, 1.0, 0.0]])
    g = s1.Grid(NN, 3.0, 0.12, F64)
    _, lab, _, _ = s1.init_trial(rows, g)
    assert lab[5, 5, 5] == 0  # equidistant: strict '>' keeps the first
    assert lab[8, 5, 5] == 1


def neumann_laplacian(label):
    """Independent sparse assembly of the label-restricted graph Laplacian
    (interior unknowns; a same-label face neighbour acts as a fixed node)."""
    import scipy.sparse as sp
    S = label.shape[0]
    n = S - 2
    idx = lambda i, j, k: ((i - 1) * n + (j - 1)) * n + (k - 1)
    rows, cols, vals = [], [], []
    for i in range(1, S - 1):
        for j in range(1, S - 1):
            for k in range(1, S - 1):
                me = idx(i, j, k)
                diag = 0
                for di, dj, dk in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                    q = (i + di, j + dj, k + dk)
                    if label[q] != label[i, j, k]:
                        continue
                    diag -= 1
                    if all(1 <= c <= S - 2 for c in q):
                        rows.append(me)
                        cols.append(idx(*q))
                        vals.append(1.0)
                rows.append(me)
                cols.append(me)
                vals.append(diag)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n ** 3, n ** 3))


def test_itp_update_equals_assembled_operator():
    NN = 7
    S = NN + 1
    rng = np.random.default_rng(1)
    g = s1.Grid(NN, NN * 0.3, 0.12, F64)
    label = np.zeros((S, S, S), np.int32)
    label[1:-1, 1:-1, 1:-1] = rng.integers(0, 3, (S - 2,) * 3)
    K = np.zeros((S, S, S))
    P = np.zeros((S, S, S))
    K[1:-1, 1:-1, 1:-1] = rng.uniform(0, 3, (S - 2,) * 3)
    P[1:-1, 1:-1, 1:-1] = rng.uniform(0, 1, (S - 2,) * 3)
    U = np.zeros((S, S, S))
    U[1:-1, 1:-1, 1:-1] = rng.standard_normal((S - 2,) * 3)
    Ln = neumann_laplacian(label)
    assert abs(Ln - Ln.T).max() == 0  # symmetric operator
    u = U[1:-1, 1:-1, 1:-1].ravel()
    diag = (K - 2 * P)[1:-1, 1:-1, 1:-1].ravel()
    expect = u + g.half_dv * (Ln @ u) + g.dtv * diag * u
    got = s1.itp_update(U, K, P, label, g)[1:-1, 1:-1, 1:-1].ravel()
    assert np.allclose(got, expect, rtol=1e-13, atol=1e-13)


def test_itp_update_sine_mode_eigenvalue():
    """Single domain, K=P=0, zero faces: sin modes are eigenvectors of the
    Dirichlet Laplacian with eigenvalue sum_a (2 cos(pi p_a / NN) - 2)."""
    NN = 10
    S = NN + 1
    g = s1.Grid(NN, 3.0, 0.12, F64)
    x = np.arange(S)
    p = (1, 2, 3)
    U = (np.sin(np.pi * p[0] * x / NN)[:, None, None] * np.sin(np.pi * p[1] * x / NN)[None, :, None]
         * np.sin(np.pi * p[2] * x / NN)[None, None, :])
    lam = sum(2 * np.cos(np.pi * q / NN) - 2 for q in p)
    label = np.zeros((S, S, S), np.int32)
    Z = np.zeros((S, S, S))
    out = s1.itp_update(U, Z, Z, label, g)
    assert np.allclose(out[1:-1, 1:-1, 1:-1], (1 + g.half_dv * lam) * U[1:-1, 1:-1, 1:-1], atol=1e-13)


def test_rc_mask_rule():
    NN = 6
    S = NN + 1
    g = s1.Grid(NN, 2.0, 0.12, F64)
    rng = np.random.default_rng(2)
    U = rng.uniform(0.5, 1, (S, S, S))
    label = np.zeros((S, S, S), np.int32)
    Z = np.zeros((S, S, S))
    mask = np.zeros((S, S, S), bool)
    mask[3, 3, 3] = True
    out = s1.itp_update(U, Z, Z, label, g, rc_mask=mask)
    assert out[3, 3, 3] == 0.0
    # a masked neighbour acts as a Neumann mirror for (2,3,3)
    nb = [U[1, 3, 3], U[2, 3, 3], U[2, 4, 3], U[2, 2, 3], U[2, 3, 4], U[2, 3, 2]]
    lap = sum(nb) - 6 * U[2, 3, 3]
    assert abs(out[2, 3, 3] - (U[2, 3, 3] + g.half_dv * lap)) < 1e-14


def test_jacobi_fixed_point_is_direct_solution():
    NN = 12
    S = NN + 1
    g = s1.Grid(NN, 3.0, 0.12, F64)
    rng = np.random.default_rng(3)
    U = np.zeros((S, S, S))
    U[1:-1, 1:-1, 1:-1] = rng.uniform(0, 1, (S - 2,) * 3)
    rhs = s1.rho_rhs(U, 2 * np.pi, g.h2, F64)
    exact = np.zeros((S, S, S))
    exact[1:-1, 1:-1, 1:-1] = spsolve_poisson(rhs)
    again = s1.smooth(exact, rhs, 0.0, 6.0, F64)
    assert np.abs(again - exact).max() < 1e-12 * np.abs(exact).max()
    # and the two-sweep Poisson step contracts toward it
    P = np.zeros((S, S, S))
    e0 = np.linalg.norm(P - exact)
    for _ in range(20):
        P = s1.jacobi(P, U, g, 2)
    assert np.linalg.norm(P - exact) < 0.7 * e0


def test_normalize_invariant_and_inactive_domains():
    NN = 12
    rows = atoms3(NN)
    rows[2, 3] = 0.0  # inactive atom: its domain keeps its scale
    g = s1.Grid(NN, 3.6, 0.12, F64)
    U, lab, K, _ = s1.init_trial(rows, g)
    rng = np.random.default_rng(4)
    lab[1:-1, 1:-1, 1:-1] = rng.integers(0, 3, (NN - 1,) * 3)
    U2, norms = s1.normalize(U, lab, rows[:, 3], g)
    after = s1.domain_norms(U2, lab, 3, g)
    assert np.allclose(after[:2], rows[:2, 3], rtol=1e-13)
    assert np.allclose(U2[lab == 2], U[lab == 2])


def test_boundary_score_bruteforce_loops():
    NN = 10
    S = NN + 1
    rows = atoms3(NN)
    g = s1.Grid(NN, 3.0, 0.12, F64)
    rng = np.random.default_rng(5)
    lab = np.zeros((S, S, S), np.int32)
    lab[1:-1, 1:-1, 1:-1] = rng.integers(0, 3, (S - 2,) * 3)
    U = rng.uniform(0, 1, (S, S, S))
    got = s1.boundary_score(U, lab, rows, g)
    h = g.h64
    ref = lab.copy()
    for i in range(1, S - 1):
        for j in range(1, S - 1):
            for k in range(1, S - 1):
                cur = lab[i, j, k]
                nbs = [lab[i + 1, j, k], lab[i - 1, j, k], lab[i, j + 1, k], lab[i, j - 1, k],
                       lab[i, j, k + 1], lab[i, j, k - 1]]
                other = sum(1 for q in nbs if q != cur)
                if other == 0:
                    continue
                scores = []
                for (ai, aj, ak, Z, rc) in rows:
                    d2 = ((i - ai) * h) ** 2 + ((j - aj) * h) ** 2 + ((k - ak) * h) ** 2 + h * h
                    scores.append(Z / math.sqrt(d2))
                best = 0.0
                bn = cur
                for n, sc in enumerate(scores):
                    if sc > best:
                        best, bn = sc, n
                if bn != cur and best > 1.1 * scores[cur] and other >= 2:
                    ref[i, j, k] = bn
    assert np.array_equal(got, ref)
    assert (got != lab).any()


def test_energy_bruteforce_loops():
    NN = 8
    S = NN + 1
    g = s1.Grid(NN, 2.4, 0.12, F64)
    rng = np.random.default_rng(6)
    lab = np.zeros((S, S, S), np.int32)
    lab[1:-1, 1:-1, 1:-1] = rng.integers(0, 2, (S - 2,) * 3)
    U = rng.uniform(0, 1, (S, S, S))
    U[0] = U[-1] = 0
    U[:, 0] = U[:, -1] = 0
    U[:, :, 0] = U[:, :, -1] = 0
    K = rng.uniform(0, 3, (S, S, S))
    P = rng.uniform(0, 1, (S, S, S))
    tot, per = s1.energy(U, K, P, lab, g, 2)
    h = g.h64
    T = VeK = Vee = 0.0
    for i in range(1, S - 1):
        for j in range(1, S - 1):
            for k in range(1, S - 1):
                u = U[i, j, k]
                L = lab[i, j, k]
                d = 0.0
                for q in ((i + 1, j, k), (i, j + 1, k), (i, j, k + 1)):
                    if lab[q] == L:
                        d += (U[q] - u) ** 2
                T += 0.5 * d * h
                VeK += -K[i, j, k] * u * u * h ** 3
                Vee += P[i, j, k] * u * u * h ** 3
    assert np.allclose(tot, [T, VeK, Vee], rtol=1e-12)
    assert np.allclose(per.sum(0), tot, rtol=1e-12)


def test_hf_force_is_minus_gradient_of_electron_nuclear_energy():
    """F_n = -d/dR_n of V_eK(R) = -sum rho K(R) h^3 at fixed density (bare atoms,
    r~ = sqrt(d^2 + h^2)): central differences of the oracle's own K."""
    NN = 12
    S = NN + 1
    rows = atoms3(NN)
    rows[:, 4] = 0.0
    g = s1.Grid(NN, 3.6, 0.12, F64)
    rng = np.random.default_rng(7)
    U = np.zeros((S, S, S))
    U[1:-1, 1:-1, 1:-1] = rng.uniform(0.1, 1, (S - 2,) * 3)
    rho = U * U
    F = s1.hf_forces(U, rows, g)
    d = 1e-4  # grid units
    for n in range(3):
        for c in range(3):
            rp, rm = rows.copy(), rows.copy()
            rp[n, c] += d
            rm[n, c] -= d
            Ep = -(rho * s1.recompute_k(rp[n:n + 1], g)).sum() * g.h3
            Em = -(rho * s1.recompute_k(rm[n:n + 1], g)).sum() * g.h3
            fd = -(Ep - Em) / (2 * d * g.h64)
            assert abs(F[n, c] - fd) < 1e-6 * (abs(fd) + 1e-3), (n, c, F[n, c], fd)


def test_nuclear_energy_and_forces_consistent():
    rows = atoms3(20)
    h = 0.2
    Z = rows[:, 3]
    F = s1.nuclear_forces(rows[:, :3], Z, h)
    assert np.allclose(F.sum(0), 0, atol=1e-12)  # Newton's third law
    d = 1e-5
    for n in range(3):
        for c in range(3):
            rp, rm = rows[:, :3].copy(), rows[:, :3].copy()
            rp[n, c] += d
            rm[n, c] -= d
            fd = -(s1.nuclear_energy(rp, Z, h) - s1.nuclear_energy(rm, Z, h)) / (2 * d * h)
            assert abs(F[n, c] - fd) < 1e-6 * (abs(fd) + 1e-6)


def test_move_nuclei_closed_form():
    pos = np.array([[50.0, 50.0, 50.0]])
    vel = np.zeros((1, 3))
    F = np.array([[0.3, -0.2, 0.0]])
    p2, v2 = s1.move_nuclei(pos, vel, F, [1.0], 0.1, 100, 0.8, 1.0, 0.98, 0.3)
    v_exp = np.clip(F[0] / 1836 * 0.8 * 0.98, -0.3, 0.3)
    assert np.allclose(v2[0], v_exp)
    assert np.allclose(p2[0], pos[0] + v_exp * 0.8 / 0.1)
    # clipping to [5, NN-5]
    p3, _ = s1.move_nuclei(np.array([[5.0, 94.9, 50.0]]), np.array([[-1.0, 1.0, 0.0]]), np.zeros((1, 3)),
                           [1.0], 0.1, 100, 0.8, 1.0, 0.98, 0.3)
    assert p3[0, 0] == 5.0 and p3[0, 1] == 95.0


def softened_hydrogen_energy(a, R=30.0, n=60000):
    """Ground state of -1/2 lap - 1/sqrt(r^2 + a^2) from the radial equation on
    a fine 1-D grid (independent of the 3-D code)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    dr = R / n
    r = dr * np.arange(1, n + 1)
    main = 1.0 / dr ** 2 - 1.0 / np.sqrt(r * r + a * a)
    off = -0.5 / dr ** 2 * np.ones(n - 1)
    H = sp.diags([off, main, off], [-1, 0, 1])
    return float(spl.eigsh(H, k=1, sigma=-1.0, which="LM")[0][0])


def test_hydrogen_energy_matches_radial_solution():
    """One bare hydrogen domain (K = 1/sqrt(r^2 + h^2), cloud/solver.py:48):
    ITP drives E = T + V_eK monotonically down to the ground state of the
    h-softened Coulomb problem, up to the O(h^2) grid error; that energy in
    turn tends to the exact -0.5 Ha as h -> 0."""
    assert abs(softened_hydrogen_energy(1e-4) + 0.5) < 1e-5
    NN, screen = 40, 12.0
    rows = np.array([[20.0, 20.0, 20.0, 1.0, 0.0]])
    st = s1.State(rows, NN, screen, F64, dv=0.12)
    st.U, st.label, st.K, _ = s1.init_trial(rows, st.g)
    st.U, _ = s1.normalize(st.U, st.label, st.Z, st.g)
    Es = []
    for step in range(1, 1201):
        st.U = s1.itp_update(st.U, st.K, st.P, st.label, st.g)
        if step % 20 == 0:
            st.U, _ = s1.normalize(st.U, st.label, st.Z, st.g)
        if step % 200 == 0:
            tot, _, _ = st.energy()
            Es.append(tot[0] + tot[1])
    assert all(b <= a + 1e-9 for a, b in zip(Es, Es[1:]))
    e_soft = softened_hydrogen_energy(screen / NN)
    assert abs(Es[-1] - e_soft) < 0.006, (Es, e_soft)
