"""Pin the multigrid oracle against mathematics (transfer-operator identities,
direct solves, DST eigenvalues) and the survey's numerical findings about
the molecule.js smoother and two-grid correction."""
import numpy as np

from oracle import mg, s1

F64 = np.float64


def test_restrict4_constants_and_affine():
    S = 17
    x = np.arange(S, dtype=F64)
    r = np.zeros((S, S, S))
    r[1:-1, 1:-1, 1:-1] = 1.0
    gc = mg.restrict4(r, F64)
    assert np.allclose(gc[1:-1, 1:-1, 1:-1], 4.0)  # weights sum to 64/64, times 4
    aff = 0.3 + 0.1 * x[:, None, None] - 0.2 * x[None, :, None] + 0.05 * x[None, None, :]
    r = np.zeros((S, S, S))
    r[1:-1, 1:-1, 1:-1] = aff[1:-1, 1:-1, 1:-1]
    gc = mg.restrict4(r, F64)
    X = 2 * np.arange(9, dtype=F64)
    ref = 0.3 + 0.1 * X[:, None, None] - 0.2 * X[None, :, None] + 0.05 * X[None, None, :]
    assert np.allclose(gc[1:-1, 1:-1, 1:-1], 4 * ref[1:-1, 1:-1, 1:-1])


def test_prolong_exact_on_affine():
    Sc = 9
    X = 2 * np.arange(Sc, dtype=F64)
    ec = 1.0 + 0.4 * X[:, None, None] + 0.1 * X[None, :, None] - 0.3 * X[None, None, :]
    f = mg.prolong(ec, F64)
    x = np.arange(2 * Sc - 1, dtype=F64)
    ref = 1.0 + 0.4 * x[:, None, None] + 0.1 * x[None, :, None] - 0.3 * x[None, None, :]
    assert np.allclose(f, ref)


def test_restriction_is_scaled_prolongation_transpose():
    """<restrict4(r), e>_coarse = (4/8) <r, P e>_fine for zero-face fields."""
    rng = np.random.default_rng(0)
    S, Sc = 17, 9
    r = np.zeros((S, S, S))
    r[1:-1, 1:-1, 1:-1] = rng.standard_normal((S - 2,) * 3)
    e = np.zeros((Sc, Sc, Sc))
    e[1:-1, 1:-1, 1:-1] = rng.standard_normal((Sc - 2,) * 3)
    lhs = (mg.restrict4(r, F64) * e).sum()
    rhs = 0.5 * (r * mg.prolong(e, F64)).sum()
    assert abs(lhs - rhs) < 1e-11 * abs(rhs)


def test_exact_solve_matches_spsolve():
    rng = np.random.default_rng(1)
    Sc = 13
    g = np.zeros((Sc, Sc, Sc))
    g[1:-1, 1:-1, 1:-1] = rng.standard_normal((Sc - 2,) * 3)
    e = mg.exact_solve(g, F64)
    ref = mg.spsolve_poisson(g[1:-1, 1:-1, 1:-1])
    assert np.allclose(e[1:-1, 1:-1, 1:-1], ref, rtol=1e-11, atol=1e-12)


def test_vcycle_contraction_and_limit():
    rng = np.random.default_rng(2)
    NN = 32
    rhs = rng.standard_normal((NN - 1,) * 3)
    exact = mg.spsolve_poisson(rhs)
    x = np.zeros((NN + 1,) * 3)
    errs = []
    for _ in range(8):
        x = mg.vcycle(x, rhs, 4, 2, 2, 1.0 / 7.0, 7.0, F64)
        errs.append(np.linalg.norm(x[1:-1, 1:-1, 1:-1] - exact) / np.linalg.norm(exact))
    rates = [b / a for a, b in zip(errs, errs[1:])]
    assert max(rates[2:]) < 0.3, rates
    assert errs[-1] < 2e-5


def test_molecule_js_smoother_is_screened():
    """The 0.3333/9 smoother's fixed point is not the exact discrete Poisson
    solution (molecule.js:747): relative difference ~1.7e-3 on 17^3, h=0.25."""
    NN = 16
    h = 0.25
    S = NN + 1
    x = (np.arange(S) - NN / 2) * h
    r2 = x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2
    rho = np.exp(-r2)
    rhs = (2 * np.pi * rho * h * h)[1:-1, 1:-1, 1:-1]
    exact = mg.spsolve_poisson(rhs)
    P = np.zeros((S, S, S))
    for _ in range(4000):
        P = s1.smooth(P, rhs, 0.3333, 9.0, F64)
    rel = np.linalg.norm(P[1:-1, 1:-1, 1:-1] - exact) / np.linalg.norm(exact)
    assert 1.0e-3 < rel < 3.0e-3, rel
    # 1/3 instead of 0.3333 converges to the exact solution
    P = np.zeros((S, S, S))
    for _ in range(4000):
        P = s1.smooth(P, rhs, 1.0 / 3.0, 9.0, F64)
    assert np.linalg.norm(P[1:-1, 1:-1, 1:-1] - exact) / np.linalg.norm(exact) < 1e-8


def test_molecule_js_twogrid_correction_is_sign_inverted():
    """The two-grid correction moves P away from the solution: its direction
    has a negative inner product with the true error (molecule.js:777-909)."""
    NN = 16
    S = NN + 1
    rng = np.random.default_rng(3)
    rhs = rng.uniform(0, 1, (S - 2,) * 3)
    exact = mg.spsolve_poisson(rhs)
    P = np.zeros((S, S, S))
    for _ in range(4):
        P = s1.smooth(P, rhs, 0.3333, 9.0, F64)
    err = exact - P[1:-1, 1:-1, 1:-1]
    P2 = mg.twogrid_s2(P, rhs, F64)
    corr = (P2 - P)[1:-1, 1:-1, 1:-1]
    assert (corr * err).sum() < 0
    # the build's correction (sign +1, exact coarse solve) reduces the error
    r = np.zeros((S, S, S))
    r[1:-1, 1:-1, 1:-1] = mg.residual(P, rhs, F64)
    ec = mg.exact_solve(mg.restrict4(r, F64), F64)
    P3 = mg.prolong_correct(P, ec, F64)
    assert np.linalg.norm(exact - P3[1:-1, 1:-1, 1:-1]) < np.linalg.norm(err)
