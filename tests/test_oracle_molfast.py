"""The mol_fast.js oracle against closed forms, invariants and brute force
(CPU only)."""
import math

import numpy as np
import pytest

from oracle import molfast as O

f32 = np.float32


def _mf(NN=16, screen=8.0, nuclei=None, **kw):
    nuc = nuclei or [O.Nucleus(6, 8, 8, 1.0), O.Nucleus(10, 8, 8, 1.0)]
    return O.MolFast(nuc, NN=NN, screen=screen, **kw)


def test_consts_closed_form():
    mf = _mf(NN=20, screen=10.0)
    c = mf.consts()
    assert c["h"] == f32(0.5) and c["h2"] == f32(0.25) and c["inv_h2"] == f32(4.0)
    assert c["dt"] == f32(0.12 * 0.25) and c["half_d"] == f32(0.06) and c["h3"] == f32(0.125)
    assert c["N2"] == 10 and c["two_pi"] == f32(2 * math.pi)
    assert c["kc"] == f32(0.5) and c["beta"] == f32(0.0) and c["R_out"] == f32(1.0)


def test_in_split_sectors():
    rng = np.random.default_rng(0)
    p = rng.normal(size=(3, 500))
    r = np.sqrt((p ** 2).sum(0))
    # hemispheres split by the sign along the axis, together covering everything once
    h0 = O.in_split(p[0], p[1], p[2], r, "hemi", 0, (0, 0, 2))
    h1 = O.in_split(p[0], p[1], p[2], r, "hemi", 1, (0, 0, 2))
    assert np.array_equal(h0, p[2] < 0) and np.all(h0 ^ h1)
    # thirds: exactly one sector each, and a 120-degree turn about the axis
    # moves a point to the next sector
    secs = np.stack([O.in_split(p[0], p[1], p[2], r, "third", s, (0, 0, 1)) for s in range(3)])
    assert np.all(secs.sum(0) == 1)
    c, s = math.cos(2 * math.pi / 3), math.sin(2 * math.pi / 3)
    q = np.stack([c * p[0] - s * p[1], s * p[0] + c * p[1], p[2]])
    sec_p, sec_q = secs.argmax(0), np.stack(
        [O.in_split(q[0], q[1], q[2], r, "third", s_, (0, 0, 1)) for s_ in range(3)]).argmax(0)
    away = np.abs(np.sin(3 * np.arctan2(p[1], p[0]))) > 1e-6  # not on a sector boundary
    assert np.all(((sec_p + 1) % 3 == sec_q)[away] | ((sec_p + 2) % 3 == sec_q)[away])
    assert len(set(((sec_q - sec_p) % 3)[away].tolist())) == 1
    # tetra: each vertex direction lies in its own sector
    for v, vec in enumerate(((1, 1, 1), (1, -1, -1), (-1, 1, -1), (-1, -1, 1))):
        x = np.array(vec, float) * 0.7
        for sidx in range(4):
            got = O.in_split(x[0], x[1], x[2], np.linalg.norm(x), "tetra", sidx, (1, 0, 0))
            assert bool(got) == (sidx == v)
    # the origin belongs to every orbital
    assert O.in_split(0.0, 0.0, 0.0, 0.0, "tetra", 3, (1, 0, 0))


def test_init_invariants_and_pointwise():
    nuc = [O.Nucleus(6, 8, 8, 3.0, rc=1.0), O.Nucleus(11, 8, 9, 1.0)]
    mf = _mf(nuclei=nuc, screen=4.0, targets=[3.0, 1.0])
    K, U, W, P = O.init_fields(mf)
    hv = mf.screen / mf.NN
    # pre-normalised to the targets (sum over every node)
    for m, t in enumerate(mf.targets):
        assert math.isclose(float(np.sum(U[m].astype(np.float64) ** 2)) * hv ** 3, t, rel_tol=1e-6)
    # partition: w in [0, 1], at most one owner per node, nothing beyond R_init
    assert W.min() >= 0 and W.max() <= 1
    assert np.all((W > 0).sum(0) <= 1)
    # pointwise restatement at sampled nodes (scalar loop, doubles)
    for (i, j, k) in [(6, 8, 8), (0, 0, 0), (11, 8, 9), (9, 7, 3), (16, 16, 16)]:
        Kk = 0.0
        for n in nuc:
            d2 = ((i - n.i) * hv) ** 2 + ((j - n.j) * hv) ** 2 + ((k - n.k) * hv) ** 2
            Kk += n.Z / math.sqrt(d2 + 4 * hv * hv)
        assert K[i, j, k] == f32(Kk)
        r0 = math.sqrt(((i - 6) * hv) ** 2 + ((j - 8) * hv) ** 2 + ((k - 8) * hv) ** 2)
        assert P[0, i, j, k] == f32(0.5 * 3.0 / (r0 + hv * hv))
    # the O cusp: U0 vanishes within rc - 3h of its own nucleus
    assert U[0, 6, 8, 8] == 0.0


def _uniform(mf, u, w, p, Kv):
    S, NE = mf.S, mf.NE
    U = np.full((NE, S, S, S), 0, np.float32)
    W = np.full((NE, S, S, S), 0, np.float32)
    P = np.full((NE, S, S, S), 0, np.float32)
    for m in range(NE):
        U[m], W[m], P[m] = f32(u[m]), f32(w[m]), f32(p[m])
    return np.full((S, S, S), f32(Kv), np.float32), U, W, P


def test_update_uniform_fields_closed_form():
    """Constant u, w, P: no Laplacian or gradient terms; the increments are
    the potential terms alone."""
    nuc = [O.Nucleus(3, 8, 8, 1.0), O.Nucleus(13, 8, 8, 1.0), O.Nucleus(8, 3, 8, 2.0)]
    mf = _mf(nuclei=nuc, targets=[1.0, 1.0, 2.0], beta_pauli=0.25)
    c = mf.consts()
    u, w, p, Kv = [0.2, 0.3, 0.5], [1.0, 0.5, 0.25], [0.1, 0.2, 0.4], 1.5
    K, U, W, P = _uniform(mf, u, w, p, Kv)
    Uo, Wo, Po = O.update(mf, K, U, W, P)
    dt = float(c["dt"])
    for m in range(3):
        others = [x for n, x in enumerate(u) if n != m]
        cm = u[m] - max(others)
        nw = w[m]  # gradients and Laplacian vanish
        assert np.allclose(Wo[m][1:-1, 1:-1, 1:-1], nw, rtol=1e-6)
        selfN = mf.targets[m]
        fm = 1.0 / selfN if selfN > 1 else 1.0
        Vother = 2 * (sum(p) - fm * p[m])
        Vpauli = 0.25 * sum(x * x for x in others)
        ii = np.arange(1, mf.S - 1)
        x_au = (ii - c["N2"]) * float(c["h"])
        # V_tilt = 0, field_x = 0: u' = u + dt (K - Vother - Vpauli) u w
        want = u[m] + dt * (Kv - Vother - Vpauli) * u[m] * w[m]
        assert np.allclose(Uo[m][1:-1, 1:-1, 1:-1], want, rtol=2e-6)
        assert np.allclose(Po[m][1:-1, 1:-1, 1:-1], p[m] + 2 * math.pi * dt * u[m] ** 2, rtol=2e-6)
        del cm, x_au
        # faces untouched
        assert Uo[m][0, 0, 0] == f32(u[m]) and Po[m][-1, 3, 3] == f32(p[m])


def test_update_laplacian_of_quadratic():
    """u_0 = k^2 (grid units), w = 1, K = P = 0, one electron: the weighted
    Laplacian is hd (u(k+1) - 2u(k) + u(k-1)) = 2 hd exactly."""
    mf = _mf(nuclei=[O.Nucleus(8, 8, 8, 1.0)])
    c = mf.consts()
    S = mf.S
    k = np.arange(S, dtype=np.float32)
    U = np.broadcast_to(k * k, (S, S, S)).astype(np.float32)[None].copy()
    W = np.ones((1, S, S, S), np.float32)
    P = np.zeros((1, S, S, S), np.float32)
    K = np.zeros((S, S, S), np.float32)
    Uo, Wo, Po = O.update(mf, K, U, W, P)
    inner = Uo[0][1:-1, 1:-1, 1:-1] - U[0][1:-1, 1:-1, 1:-1]
    # FP32 spacing of u (up to 256) bounds the error of the difference
    assert np.allclose(inner, 2 * float(c["half_d"]), rtol=0, atol=2e-5)
    # the advection of w is driven by cm = u (no other electron): w stays 1 (clamped)
    assert np.all(Wo[0] == 1.0)


def test_update_poisson_part_is_linear_in_the_source():
    mf = _mf()
    rng = np.random.default_rng(1)
    S = mf.S
    K = np.zeros((S, S, S), np.float32)
    U = rng.uniform(0, 1, (2, S, S, S)).astype(np.float32)
    W = np.zeros((2, S, S, S), np.float32)
    P0 = np.zeros((2, S, S, S), np.float32)
    _, _, Pa = O.update(mf, K, U, W, P0)
    _, _, Pb = O.update(mf, K, (U * f32(2.0)).astype(np.float32), W, P0)
    assert np.allclose(Pb, 4 * Pa, rtol=1e-6)


def test_reduce_closed_forms():
    mf = _mf(nuclei=[O.Nucleus(8, 8, 8, 1.0)])
    c = mf.consts()
    S, n = mf.S, mf.S - 2
    a = 0.01
    kk = np.arange(S, dtype=np.float32)
    U = (np.broadcast_to(f32(a) * kk, (S, S, S)).astype(np.float32))[None].copy()
    W = np.ones((1, S, S, S), np.float32)
    P = np.zeros((1, S, S, S), np.float32)
    K = np.zeros((S, S, S), np.float32)
    s = O.reduce_sums(mf, K, U, W, P)
    h3 = float(c["h3"])
    inner = np.arange(1, S - 1)
    assert math.isclose(s[0], n * n * float(np.sum((a * inner) ** 2)) * h3, rel_tol=1e-5)
    # T: forward difference along k only, = kc a^2 h per interior node
    assert math.isclose(s[1], n ** 3 * 0.5 * a * a * float(c["h"]), rel_tol=1e-4)
    assert s[2] == 0.0 and s[3] == 0.0
    # dipole of a density symmetric about the centre in i and j vanishes
    assert abs(s[4]) < 1e-12 and abs(s[5]) < 1e-12


def test_twogrid_pieces():
    mf = _mf()
    S, NC = mf.S, mf.NN // 2
    fine = np.full((S, S, S), f32(3.0), np.float32)
    co = O.restrict_fw(mf, fine)
    assert np.all(co[1:NC, 1:NC, 1:NC] == 3.0) and co[0].max() == 0
    Ec = np.zeros((NC + 1,) * 3, np.float32)
    Ec[:] = f32(2.0)
    Pm = np.zeros((S, S, S), np.float32)
    out = O.prolong_correct(mf, Pm, Ec)
    assert np.all(out[1:-1, 1:-1, 1:-1] == 1.0) and out[0].max() == 0
    # a zero coarse rhs keeps a zero error at zero; a uniform interior error decays by 0.3333 + 6/9 < 1
    E = np.zeros((NC + 1,) * 3, np.float32)
    assert np.all(O.coarse_smooth(mf, E, E) == 0)


def test_forces_linear_potential():
    nuc = [O.Nucleus(8, 8, 8, 2.0), O.Nucleus(5, 8, 8, 0.0)]
    mf = _mf(nuclei=nuc)
    c = mf.consts()
    S = mf.S
    alpha = 0.125
    ii = np.arange(S, dtype=np.float32)
    P = np.zeros((2, S, S, S), np.float32)
    P[0] = np.broadcast_to((f32(alpha) * ii)[:, None, None], (S, S, S))
    F = O.forces_elec(mf, P)
    # grad_i P = alpha / h exactly on the grid; F = 2 Z <grad P>
    assert math.isclose(float(F[0, 0]), 2 * 2.0 * alpha * float(c["inv_h"]), rel_tol=1e-5)
    assert F[0, 1] == 0 and F[0, 2] == 0
    assert np.all(F[1] == 0)  # Z = 0 feels no force


def test_recompute_k_matches_initial_k():
    mf = _mf(nuclei=[O.Nucleus(7, 8, 9, 1.0)])
    K0, _, _, _ = O.init_fields(mf)
    K1 = O.recompute_k(mf)
    assert np.allclose(K0, K1, rtol=2e-6)


def test_extract_scalar_loop():
    mf = _mf(nuclei=[O.Nucleus(6, 8, 8, 1.0), O.Nucleus(10, 8, 8, 1.0)])
    K, U, W, P = O.init_fields(mf)
    out = O.extract(mf, U, W, P)
    S, NE, N2 = mf.S, mf.NE, mf.NN // 2
    for m in range(NE):
        for i in (0, 3, 8, 16):
            for j in (0, 5, 16):
                v = U[m, i, j, N2] if W[m, i, j, N2] > f32(0.1) else 0.0
                assert out[m * S * S + i * S + j] == v
            assert out[NE * S * S + m * S + i] == U[m, i, N2, N2]
            assert out[2 * NE * S * S + m * S + i] == W[m, i, N2, N2]
    for i in (0, 7, 16):
        assert out[NE * S * S + NE * S + i] == f32(2.0) * (P[0, i, N2, N2] + P[1, i, N2, N2])


def test_do_steps_schedule(monkeypatch):
    mf = _mf(vcycle_interval=10)
    K, U, W, P = O.init_fields(mf)
    calls = []
    real_norm, real_tg = O.normalize, O.twogrid
    monkeypatch.setattr(O, "normalize", lambda *a: (calls.append("n"), real_norm(*a))[1])
    monkeypatch.setattr(O, "twogrid", lambda *a: (calls.append("v"), real_tg(*a))[1])
    U2, W2, P2, sums = O.do_steps(mf, K, U, W, P, 25)
    # normalisation at steps 20 and 25 (last), 2-grid at 10 and 20 (after the norm at 20)
    assert calls == ["v", "n", "v", "n"]
    assert np.all(np.isfinite(U2)) and np.all(np.isfinite(sums))
    for m in range(mf.NE):
        # the last normalisation used the sums of that same step
        norm = float(np.sum(U2[m][1:-1, 1:-1, 1:-1].astype(np.float64) ** 2)) * float(mf.consts()["h3"])
        assert math.isclose(norm, mf.targets[m], rel_tol=1e-5)


@pytest.mark.parametrize("shell", [False, True])
def test_split_orbitals_update_keeps_sectors(shell):
    """Shell mode zeroes w outside an orbital's enforced sector."""
    nuc = [O.Nucleus(8, 8, 8, 2.0, split="hemi", split_idx=0, split_axis=(1, 0, 0)),
           O.Nucleus(8, 8, 8, 2.0, split="hemi", split_idx=1, split_axis=(1, 0, 0))]
    mf = _mf(nuclei=nuc, shell_mode=shell)
    tab = mf.orbital_tables()
    assert tab[1]["group"] == 0  # split siblings share a group
    K, U, W, P = O.init_fields(mf)
    Uo, Wo, Po = O.update(mf, K, U, W, P)
    if shell:
        assert np.all(Wo[0][9:, :, :] == 0) and np.all(Wo[1][1:8, 1:-1, 1:-1] == 0)
