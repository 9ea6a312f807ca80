"""Pins of the S2 oracle's per-atom fields (Z_nuc, initZeff, initRcut,
angular splits), its K refresh, the unique-nucleus E_KK and the direct
Pother scheme (CPU only), by closed forms, symmetry and invariants."""
import math

import numpy as np
import pytest

from oracle import s1, s2

F32 = np.float32


def _grid(NN=16, h=0.3):
    return s1.Grid(NN, NN * h, 0.12, F32)


def test_in_split_hemi_is_the_sign_of_the_axis_projection():
    rng = np.random.default_rng(0)
    d = rng.normal(size=(3, 500)).astype(F32)
    axis = (0.3, -0.5, 0.8)
    a = np.array(axis) / np.linalg.norm(axis)
    proj = a @ d.astype(np.float64)
    keep = np.abs(proj) > 1e-4
    r = np.sqrt((d.astype(np.float64) ** 2).sum(0)).astype(F32)
    for idx, want in ((0, proj < 0), (1, proj >= 0)):
        e = s2.resolve_ext([[0, 0, 0, 1, 0]], [{"split": "hemi", "split_idx": idx, "split_axis": axis}])[0]
        got = s2.in_split(d[0], d[1], d[2], r, e, F32)
        assert np.array_equal(got[keep], want[keep])


def test_in_split_third_rotates_to_the_next_sector():
    """A point rotated by +120 degrees about the axis lies in the next third."""
    rng = np.random.default_rng(1)
    axis = np.array([0.2, 0.1, 0.97])
    axis /= np.linalg.norm(axis)
    pts = rng.normal(size=(400, 3))

    def rot(v, ang):
        k = axis
        return v * math.cos(ang) + np.cross(k, v) * math.sin(ang) + np.outer(v @ k, k) * (1 - math.cos(ang))

    def sector(v):
        out = np.full(len(v), -1)
        r = np.sqrt((v ** 2).sum(1)).astype(F32)
        for idx in range(3):
            e = s2.resolve_ext([[0, 0, 0, 1, 0]], [{"split": "third", "split_idx": idx, "split_axis": axis,
                                                    "split_rot": 0.4}])[0]
            inside = s2.in_split(v[:, 0].astype(F32), v[:, 1].astype(F32), v[:, 2].astype(F32), r, e, F32)
            out[inside] = idx
        return out

    s0 = sector(pts)
    s1_ = sector(rot(pts, 2 * math.pi / 3))
    # the sector advances by one (all but points within rounding of an edge)
    assert (s0 >= 0).all() and (s1_ >= 0).all()
    agree = (s1_ - s0) % 3
    assert (agree == 1).mean() > 0.99


def test_bare_nucleus_adds_to_k_but_owns_no_domain():
    g = _grid()
    rows = np.array([[5.0, 8.0, 8.0, 1.0, 0.0], [11.0, 8.0, 8.0, 0.0, 0.0]])
    U, lab, K, _ = s2.init_s2(rows, g, [{}, {"Z_nuc": 1.0}])
    rows_both = rows.copy()
    rows_both[1, 3] = 1.0
    _, lab_b, K_b, _ = s2.init_s2(rows_both, g)
    assert np.array_equal(K, K_b)
    assert (lab == 0).all() and (lab_b == 1).any()


def test_init_zeff_and_rcut_shape_the_trial():
    g = _grid()
    rows = np.array([[8.0, 8.0, 8.0, 2.0, 0.0]])
    zeff, rcut = 1.3, 1.0
    U, lab, K, _ = s2.init_s2(rows, g, [{"init_zeff": zeff, "init_rcut": rcut}])
    x = np.arange(g.S) * g.h64
    d = np.sqrt(((x[:, None, None] - 8 * g.h64) ** 2 + (x[None, :, None] - 8 * g.h64) ** 2
                 + (x[None, None, :] - 8 * g.h64) ** 2))
    r_soft = np.maximum(d, 2 * g.h64)
    closed = zeff ** 2 / math.sqrt(math.pi) * np.exp(-zeff * r_soft)
    inside = d <= rcut * (1 - 1e-6)
    assert np.allclose(U[inside], closed[inside], rtol=2e-6)
    assert (U[d > rcut * (1 + 1e-6)] == 0).all()
    # K still uses the nuclear charge Z
    assert np.allclose(K, 2.0 / r_soft, rtol=2e-6)


def test_hemi_pair_partitions_by_the_plane():
    g = _grid()
    c = 8.0
    rows = np.array([[c, c, c, 1.0, 0.0], [c, c, c, 1.0, 0.0]])
    ext = [{"split": "hemi", "split_idx": 0, "split_axis": (1, 0, 0)},
           {"split": "hemi", "split_idx": 1, "split_axis": (1, 0, 0)}]
    U, lab, K, _ = s2.init_s2(rows, g, ext)
    i = np.arange(g.S)[:, None, None] * np.ones((1, g.S, g.S))
    assert (lab[i < c] == 0).all() and (lab[i > c] == 1).all()


def test_recompute_k_s2_closed_form_and_init_agreement():
    g = _grid()
    rows = np.array([[7.0, 9.0, 8.0, 1.0, 0.0], [10.0, 6.0, 8.0, 2.0, 0.5]])
    K = s2.recompute_k_s2(rows, g, [{"Z_nuc": 1.5}, {}])
    x = np.arange(g.S)
    ref = np.zeros((g.S,) * 3)
    for (ai, aj, ak, Z, rc), zn in zip(rows, (1.5, 2.0)):
        r = np.sqrt(((x[:, None, None] - ai) * g.h64) ** 2 + ((x[None, :, None] - aj) * g.h64) ** 2
                    + ((x[None, None, :] - ak) * g.h64) ** 2)
        ref += zn / (np.maximum(r, rc) if rc > 0 else np.maximum(r, 2 * g.h64))
    assert np.allclose(K, ref, rtol=3e-6)
    # rc = 0 atoms: the refresh equals the init K (same clamp, same charge;
    # offsets formed in another order: FP32 rounding)
    K_init = s2.init_s2(rows[:1], g, [{"Z_nuc": 1.5}])[2]
    assert np.allclose(s2.recompute_k_s2(rows[:1], g, [{"Z_nuc": 1.5}]), K_init, rtol=1e-6)


def test_nuclear_energy_unique_merges_colocated_nuclei():
    h = 0.2
    pos = [(10, 10, 10), (10, 10, 10), (14, 10, 10)]
    Z = [1.0, 1.0, 2.0]
    e = s2.nuclear_energy_unique(pos, Z, None, h)
    d = math.sqrt((4 * h) ** 2 + 0.04 * h * h)
    assert e == pytest.approx(2.0 * 2.0 / d, rel=1e-14)
    # a bare nucleus counts with its Z_nuc; an empty atom not at all
    e2 = s2.nuclear_energy_unique([(0, 0, 0), (5, 0, 0), (9, 9, 9)], [1.0, 0.0, 0.0], [0.0, 3.0, 0.0], h)
    assert e2 == pytest.approx(3.0 / math.sqrt((5 * h) ** 2 + 0.04 * h * h), rel=1e-14)


def test_pdirect_init_scalar_loop_and_sweeps_step():
    g = _grid(12, 0.4)
    rows = np.array([[4.3, 6.0, 6.0, 1.0, 0.0], [7.6, 6.2, 5.5, 2.0, 0.4], [6.0, 8.9, 6.0, 0.0, 0.0]])
    P0 = s2.pdirect_init(rows, g, 0)
    for (i, j, k) in ((1, 1, 1), (5, 6, 7), (10, 3, 9)):
        v = 0.0
        for n in (1,):  # atom 2 has Z = 0, atom 0 is the domain itself
            ai, aj, ak = np.floor(rows[n, :3])
            v += 0.5 / math.sqrt(((i - ai) * g.h64) ** 2 + ((j - aj) * g.h64) ** 2 + ((k - ak) * g.h64) ** 2
                                 + g.h64 ** 2)
        assert P0[i, j, k] == pytest.approx(v, rel=2e-6)
    assert (P0[0] == 0).all()
    Pdom, Pother = s2.direct_sweeps_setup(rows, g)
    assert np.allclose(Pother, Pdom[0] + Pdom[1], rtol=1e-6)
    U, lab, K, _ = s2.init_s2(rows, g)
    Pdom2, Pother2 = s2.direct_sweeps_step(U, lab, Pdom, Pother, rows[:, 3], g)
    I = (slice(1, -1),) * 3
    for m in (0, 1):
        sel = lab[I] == m
        assert np.array_equal(Pother2[I][sel], Pdom2[m][I][sel])
    # the sweeps solve nabla^2 P_m = -2 pi rho_other: with many sweeps from a
    # zero start P_m is linear in rho_other
    Z0 = [np.zeros_like(P) for P in Pdom]
    A, _ = s2.direct_sweeps_step(U, lab, Z0, Pother, rows[:, 3], g, sweeps=40)
    B, _ = s2.direct_sweeps_step(U * F32(math.sqrt(2.0)), lab, Z0, Pother, rows[:, 3], g, sweeps=40)
    assert np.allclose(B[0], 2 * A[0], rtol=1e-5, atol=1e-9)
