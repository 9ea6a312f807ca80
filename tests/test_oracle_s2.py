"""Pin the S2 oracle (molecule.js restatement) against explicit-loop
restatements, closed forms and invariants."""
import math

import numpy as np

from oracle import s1, s2

F64 = np.float64


def atoms3(NN):
    c = NN / 2
    return np.array([[c - 2.3, c + 0.7, c + 0.2, 1.0, 0.0],
                     [c + 0.4, c - 0.3, c - 0.6, 2.0, 0.4],
                     [c + 2.1, c + 1.9, c + 0.9, 1.0, 0.0]])


def test_init_s2_bruteforce_all_nodes():
    NN, screen = 10, 3.0
    rows = atoms3(NN)
    g = s1.Grid(NN, screen, 0.12, F64)
    U, lab, K, _ = s2.init_s2(rows, g)
    h = screen / NN
    for (i, j, k) in [(0, 0, 0), (0, 5, 7), (5, 5, 5), (10, 3, 4), (4, 6, 5)]:
        Kv, best, bn = 0.0, 0.0, 0
        for n, (ai, aj, ak, Z, rc) in enumerate(rows):
            ai, aj, ak = math.floor(ai), math.floor(aj), math.floor(ak)
            r2 = ((i - ai) * h) ** 2 + ((j - aj) * h) ** 2 + ((k - ak) * h) ** 2
            r = max(math.sqrt(r2 + 0.04 * h * h), rc) if rc > 0 else max(math.sqrt(r2), 2 * h)
            Kv += Z / r
            u = Z * Z * 0.5641895835 * math.exp(-Z * r)
            if u > best:
                best, bn = u, n
        assert abs(K[i, j, k] - Kv) < 1e-12 * Kv
        assert abs(U[i, j, k] - best) < 1e-12 * best
        assert lab[i, j, k] == bn


def test_rc_mask_bruteforce():
    NN = 10
    rows = atoms3(NN)
    rows[1, 4] = 0.7
    g = s1.Grid(NN, 3.0, 0.12, F64)
    _, lab, _, _ = s2.init_s2(rows, g)
    m = s2.rc_mask(lab, rows, g)
    h = 0.3
    for idx in np.ndindex(*lab.shape):
        a = rows[lab[idx]]
        r = math.sqrt(sum(((idx[c] - math.floor(a[c])) * h) ** 2 for c in range(3)))
        assert m[idx] == (a[4] > 0 and r < a[4])
    assert m.any()


def levelset_loops(U, lab, W, dt_w, curv, lock=None):
    S = U.shape[0]
    Lo, Wo = lab.copy(), W.copy()
    for i in range(1, S - 1):
        for j in range(1, S - 1):
            for k in range(1, S - 1):
                me = lab[i, j, k]
                cnt, br, bl = 0, 0.0, me
                for q in ((i + 1, j, k), (i - 1, j, k), (i, j + 1, k), (i, j - 1, k), (i, j, k + 1), (i, j, k - 1)):
                    if lab[q] == me:
                        cnt += 1
                    elif U[q] ** 2 > br:
                        br, bl = U[q] ** 2, lab[q]
                w = W[i, j, k]
                if cnt == 6:
                    Wo[i, j, k] = max(w, 1.0)
                    continue
                if lock is not None and lock[me] > 0 and lock[me] == lock[bl]:
                    Wo[i, j, k] = max(w, 0.5)
                    continue
                rho = U[i, j, k] ** 2
                v = 0.0 if rho + br < 1e-20 else (rho - br) / (rho + br)
                w = w + dt_w * v + curv * (cnt - 3) / 3.0
                if w < 0:
                    Lo[i, j, k] = bl
                    w = 0.1
                Wo[i, j, k] = w
    return Lo, Wo


def test_levelset_matches_loops_and_invariants():
    NN = 9
    S = NN + 1
    rng = np.random.default_rng(0)
    g = s1.Grid(NN, 2.7, 0.12, F64)
    lab = np.zeros((S, S, S), np.int32)
    lab[1:-1, 1:-1, 1:-1] = (np.arange(S - 2)[:, None, None] >= 4).astype(np.int32) + \
        (rng.random((S - 2,) * 3) < 0.15).astype(np.int32)
    U = rng.uniform(0, 1, (S, S, S))
    W = rng.uniform(-0.2, 1.2, (S, S, S))
    for lock in (None, np.array([1, 1, 0])):
        L1, W1 = s2.levelset(U, lab, W, g, 0.5, 0.15, lock)
        L2, W2 = levelset_loops(U, lab, W, 0.5, 0.15, lock)
        assert np.array_equal(L1, L2)
        assert np.allclose(W1, W2, atol=1e-14)
    L1, W1 = s2.levelset(U, lab, W, g, 0.5, 0.15)
    flipped = L1 != lab
    assert flipped.any()
    # a flip goes to a label present among the six neighbours
    for (i, j, k) in np.argwhere(flipped):
        nb = {lab[i + 1, j, k], lab[i - 1, j, k], lab[i, j + 1, k], lab[i, j - 1, k], lab[i, j, k + 1], lab[i, j, k - 1]}
        assert L1[i, j, k] in nb
    inner = np.ones_like(lab, bool)
    for axis in range(3):
        for sign in (1, -1):
            inner[1:-1, 1:-1, 1:-1] &= s1._nb(lab, axis, sign) == lab[1:-1, 1:-1, 1:-1]
    inner[0], inner[-1], inner[:, 0], inner[:, -1], inner[:, :, 0], inner[:, :, -1] = [False] * 6
    assert (W1[inner] >= 1.0).all()


def energy_loops(U, K, P, lab, rows, h):
    S = U.shape[0]
    n = len(rows)
    per = np.zeros((n, 3))
    dip = np.zeros(3)
    fl = np.floor(rows[:, :3])

    def excl(i, j, k, L):
        r = math.sqrt(((i - fl[L, 0]) * h) ** 2 + ((j - fl[L, 1]) * h) ** 2 + ((k - fl[L, 2]) * h) ** 2)
        return r < max(rows[L, 4], 2 * h)

    for i in range(1, S - 1):
        for j in range(1, S - 1):
            for k in range(1, S - 1):
                L = lab[i, j, k]
                if excl(i, j, k, L):
                    continue
                u = U[i, j, k]
                t = 0.0
                for q in ((i + 1, j, k), (i, j + 1, k), (i, j, k + 1)):
                    if lab[q] == L and not excl(*q, L):
                        t += (U[q] - u) ** 2
                per[L] += [0.5 * t * h, -K[i, j, k] * u * u * h ** 3, P[i, j, k] * u * u * h ** 3]
                dip += u * u * h ** 3 * np.array([i * h, j * h, k * h])
    return per, dip


def test_energy_s2_matches_loops():
    NN = 10
    S = NN + 1
    rows = atoms3(NN)
    g = s1.Grid(NN, 3.0, 0.12, F64)
    rng = np.random.default_rng(1)
    U, lab, K, _ = s2.init_s2(rows, g)
    P = rng.uniform(0, 1, (S, S, S))
    tot, per, dip = s2.energy_s2(U, K, P, lab, rows, g)
    per2, dip2 = energy_loops(U, K, P, lab, rows, 0.3)
    assert np.allclose(per, per2, rtol=1e-12, atol=1e-15)
    assert np.allclose(dip, dip2, rtol=1e-12)


def test_dipole_of_symmetric_density():
    NN = 12
    S = NN + 1
    g = s1.Grid(NN, 3.6, 0.12, F64)
    x = np.arange(S) - 6.0
    U = np.exp(-0.3 * (x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2))
    lab = np.zeros((S, S, S), np.int32)
    rows = np.array([[6.0, 6.0, 6.0, 1.0, 0.0]])
    _, _, dip = s2.energy_s2(U, np.zeros_like(U), np.zeros_like(U), lab, rows, g)
    # kept density is symmetric about node 6 (the exclusion ball is centred there)
    I = (slice(1, -1),) * 3
    keep = np.ones((S - 2,) * 3, bool)
    ii = np.arange(1, S - 1) - 6.0
    keep &= np.sqrt((ii[:, None, None] * 0.3) ** 2 + (ii[None, :, None] * 0.3) ** 2 + (ii[None, None, :] * 0.3) ** 2) >= 0.6
    mass = (U[I] ** 2 * 0.3 ** 3)[keep].sum()
    assert np.allclose(dip, 6 * 0.3 * mass, rtol=1e-12)


def test_gradp_linear_potential_closed_form():
    NN = 24
    S = NN + 1
    g = s1.Grid(NN, 4.8, 0.12, F64)
    x = np.arange(S) * 0.2
    a, b, c = 0.3, -0.7, 1.1
    P = a * x[:, None, None] + b * x[None, :, None] + c * x[None, None, :]
    rows = np.array([[12.4, 11.7, 12.2, 2.0, 0.0], [6.0, 6.0, 6.0, 0.0, 0.0]])
    F = s2.gradp_forces(P, rows, g, R=5)
    assert np.allclose(F[0], [2 * 2 * a, 2 * 2 * b, 2 * 2 * c], rtol=1e-12)
    assert not F[1].any()


def test_pother_direct_splits_the_total_potential():
    """Two domains: the V-cycle is linear in its right-hand side, so from zero
    P_0 + P_1 (each solved with the other domain's charge) equals one
    V-cycle on the total charge; the gathered P holds P_1 on domain 0 and P_0
    on domain 1."""
    from oracle import mg
    NN, screen = 16, 16 * 0.4
    g = s1.Grid(NN, screen, 0.12, np.float64)
    S = NN + 1
    rng = np.random.default_rng(4)
    U = np.zeros((S, S, S))
    U[1:-1, 1:-1, 1:-1] = rng.uniform(0.1, 1.0, (S - 2,) * 3)
    label = np.zeros((S, S, S), np.int32)
    label[S // 2:] = 1
    Pdom = [np.zeros((S, S, S)), np.zeros((S, S, S))]
    P, Pd = s2.pother_direct(U, label, Pdom, [1.0, 1.0], g, 3, 2, 2, 1 / 7, 7.0)
    tot = mg.vcycle(np.zeros((S, S, S)), s1.rho_rhs(U, g.two_pi, g.h2, np.float64), 3, 2, 2, 1 / 7, 7.0, np.float64)
    assert np.allclose(Pd[0] + Pd[1], tot, rtol=1e-12, atol=1e-12)
    I = (slice(1, -1),) * 3
    li = label[I]
    assert np.array_equal(P[I][li == 0], Pd[0][I][li == 0])
    assert np.array_equal(P[I][li == 1], Pd[1][I][li == 1])
    # domain 0's potential comes from domain 1's charge: larger there on average
    assert Pd[0][I][li == 1].mean() > Pd[0][I][li == 0].mean()


def test_display_slice_matches_scalar_loops():
    """The vectorised slice against a per-pixel restatement of the WGSL body."""
    NN = 10
    S = NN + 1
    rng = np.random.default_rng(8)
    U = rng.uniform(0.0, 0.01, (S, S, S)).astype(np.float32)
    U[:, :, 5][rng.random((S, S)) < 0.3] = 0.0
    label = rng.integers(0, 3, (S, S, S)).astype(np.int32)
    K = rng.uniform(0, 5, (S, S, S)).astype(np.float32)
    atoms = np.array([[3, 3, 3, 1.0, 0], [6, 6, 6, 2.0, 0.4], [2, 8, 5, 4.0, 0.3]])
    sk = 5
    got = s2.display_slice(U, label, K, atoms, sk, NN)
    f = np.float32
    ref = np.zeros(3 * S * S + 2 * S, f)
    for i in range(S):
        ref[3 * S * S + i] = K[i, sk, sk]
        ref[3 * S * S + S + i] = 1.0
        for j in range(S):
            if i < 1 or i >= NN or j < 1 or j >= NN:
                continue
            u = f(U[i, j, sk])
            lbl = int(label[i, j, sk])
            ref[i * S + j] = u * u
            ref[S * S + i * S + j] = f(lbl) + f(atoms[lbl, 3]) * f(1000.0)
            dens = u * u
            bnd = 0.0
            if 1 < i < NN - 1 and lbl != label[i + 1, j, sk]:
                nd = f(U[i + 1, j, sk]) * f(U[i + 1, j, sk])
                bnd = 1.0 if (dens > f(1e-6) or nd > f(1e-6)) else 2.0
            if bnd < 1.5 and 1 < j < NN - 1 and lbl != label[i, j + 1, sk]:
                nd = f(U[i, j + 1, sk]) * f(U[i, j + 1, sk])
                bnd = max(bnd, 1.0) if (dens > f(1e-6) or nd > f(1e-6)) else 2.0
            ref[2 * S * S + i * S + j] = bnd
    assert np.array_equal(got, ref)
    assert (got[2 * S * S:3 * S * S] == 2).any() and (got[2 * S * S:3 * S * S] == 1).any()


def _sic_case(NN=16):
    from oracle import s1
    rows = np.array([[5.3, 8.2, 8.1, 1.0, 0.0], [10.6, 8.4, 7.9, 2.0, 0.0]])
    g = s1.Grid(NN, NN * 0.3, 0.12, np.float32)
    U, lab, K, _ = s1.init_trial(rows, g)
    U, _ = s1.normalize(U, lab, rows[:, 3], g)
    P = np.random.default_rng(3).uniform(0, 1, U.shape).astype(np.float32)
    return rows, g, U, lab, P


def test_self_potential_linear_and_additive():
    from oracle import s2
    rows, g, U, lab, P = _sic_case()
    a, c = 0.3333, 9.0
    x0 = s2.self_potential(U, lab, 0, g, a, c)
    x1 = s2.self_potential(U, lab, 1, g, a, c)
    # linear in the density: U -> sqrt(2) U doubles the self-potential
    x0b = s2.self_potential((U * np.float32(np.sqrt(2.0))).astype(np.float32), lab, 0, g, a, c)
    assert np.allclose(x0b, 2 * x0, rtol=1e-5, atol=1e-7)
    # additive over domains: the solve of the total density (every node one domain)
    allone = np.where(lab >= 0, 0, 0).astype(lab.dtype)
    xt = s2.self_potential(U, allone, 0, g, a, c)
    assert np.allclose(x0 + x1, xt, rtol=1e-5, atol=1e-6)
    assert x0.max() > 0 and x1.max() > 0


def test_pother_frame_schedule_and_mask():
    from oracle import s2
    rows, g, U, lab, P = _sic_case()
    Z = rows[:, 3]
    assert np.array_equal(s2.pother_frame(U, lab, P, Z, g, 0, 1, 0.3333, 9.0), P)  # frame 0: no SIC
    assert np.array_equal(s2.pother_frame(U, lab, P, Z, g, 3, 2, 0.3333, 9.0), P)  # not a SIC frame
    out = s2.pother_frame(U, lab, P, Z, g, 4, 2, 0.3333, 9.0)
    I = (slice(1, -1),) * 3
    x1 = s2.self_potential(U, lab, 1, g, 0.3333, 9.0)
    on1 = (lab[I] & 0x7FFF) == 1
    # domain 1 (Z = 2) loses half its self-potential on its own nodes only
    d = (P - out)[I]
    assert np.allclose(d[on1], 0.5 * x1[I][on1], rtol=1e-6, atol=1e-7)
    assert np.array_equal(out[0], P[0])  # faces untouched
