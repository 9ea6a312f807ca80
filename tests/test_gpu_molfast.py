"""The mol_fast.js fused per-electron solver on the GPU (librq.so through the
C ABI) against the oracle on the same inputs."""
import math

import numpy as np
import pytest

from oracle import molfast as O
from realmolecule_amd.molfast import MolFast
from tests.helpers import ulp_diff

pytestmark = pytest.mark.gpu

f32 = np.float32


def _pair(NN, nuclei, targets=None, **user):
    """An oracle run description and the matching MolFast host."""
    cfg = {"USER_NN": NN, "USER_SCREEN": user.pop("screen", 8.0), "USER_NUCLEI": nuclei}
    if targets is not None:
        cfg["USER_NORM_TARGETS"] = targets
    okw = {}
    for key, name in (("USER_SHELL_MODE", "shell_mode"), ("USER_BETA_PAULI", "beta_pauli"),
                      ("USER_ALPHA_TF", "alpha_tf"), ("USER_FULL_SELF", "full_self"), ("USER_FIELD_X", "field_x"),
                      ("USER_KC_MULTI", "kc_multi"), ("USER_VCYCLE_INTERVAL", "vcycle_interval"),
                      ("USER_COARSE_ITERS", "coarse_iters")):
        if key in user:
            cfg[key] = user[key]
            okw[name] = user[key]
    nuc = [O.Nucleus(n["i"], n["j"], n["k"], n["Z"], rc=n.get("rc", 0.0), r_out=n.get("r_out", 0.0),
                     tilt_y=n.get("tilt_y", 0.0), tilt_x=n.get("tilt_x", 0.0), hs=n.get("hs", 0.0),
                     split=n.get("split", "sphere"), split_idx=n.get("split_idx", 0),
                     split_axis=tuple(n.get("split_axis", (1, 0, 0)))) for n in nuclei]
    mf = O.MolFast(nuc, NN=NN, screen=cfg["USER_SCREEN"], targets=targets, **okw)
    return mf, MolFast(cfg)


def _fields(host, NE):
    U = np.stack([host.field("U", m) for m in range(NE)])
    W = np.stack([host.field("W", m) for m in range(NE)])
    P = np.stack([host.field("P", m) for m in range(NE)])
    return host.field("K"), U, W, P


def _load(host, K, U, W, P):
    host.set_field("K", 0, K)
    for m in range(U.shape[0]):
        host.set_field("U", m, U[m])
        host.set_field("W", m, W[m])
        host.set_field("P", m, P[m])


WATER = [dict(i=12, j=12, k=12, Z=3, rc=0.6), dict(i=15, j=14, k=12, Z=1), dict(i=9, j=14, k=12, Z=1)]


def test_init_matches_oracle():
    mf, host = _pair(24, WATER, targets=[3, 1, 1])
    with host:
        K, U, W, P = _fields(host, 3)
        Ko, Uo, Wo, Po = O.init_fields(mf)
        assert ulp_diff(K, Ko) == 0
        assert ulp_diff(P, Po) == 0
        assert np.array_equal(W, Wo)
        # exp() and the FP64 norm sum may differ in the last double bit
        assert ulp_diff(U, Uo) <= 1


@pytest.mark.parametrize("case", ["plain", "shell_hemi", "options"])
def test_update_step_matches_oracle(case):
    """One doSteps(1): fused update, reduction and normalisation, from the
    oracle's initial fields."""
    if case == "plain":
        mf, host = _pair(24, WATER, targets=[3, 1, 1])
    elif case == "shell_hemi":
        nuc = [dict(i=12, j=12, k=12, Z=2, rc=0.5, split="hemi", split_idx=0, split_axis=(1, 1, 0)),
               dict(i=12, j=12, k=12, Z=2, rc=0.5, split="hemi", split_idx=1, split_axis=(1, 1, 0), r_out=2.0),
               dict(i=16, j=12, k=13, Z=1, tilt_y=0.3)]
        mf, host = _pair(24, nuc, targets=[1, 1, 1], USER_SHELL_MODE=True, USER_BETA_PAULI=0.5)
    else:
        nuc = [dict(i=12, j=12, k=12, Z=3, rc=0.6, tilt_x=0.2), dict(i=15, j=14, k=12, Z=1),
               dict(i=9, j=14, k=12, Z=1)]
        mf, host = _pair(24, nuc, targets=[2, 1, 1], USER_FULL_SELF=True, USER_FIELD_X=0.05, USER_KC_MULTI=1.5)
    with host:
        Ko, Uo, Wo, Po = O.init_fields(mf)
        _load(host, Ko, Uo, Wo, Po)
        host.do_steps(1)
        K, U, W, P = _fields(host, mf.NE)
        U1, W1, P1 = O.update(mf, Ko, Uo, Wo, Po)
        sums = O.reduce_sums(mf, Ko, U1, W1, P1)
        U1 = O.normalize(mf, U1, sums)
        assert ulp_diff(W, W1) == 0
        assert ulp_diff(P, P1) == 0
        assert ulp_diff(U, U1) <= 1
        assert np.allclose(host.sums, sums, rtol=1e-11, atol=1e-14)


def test_update_sector_splits_and_tf():
    """'third' / 'tetra' sectors (atan2f) and the Thomas-Fermi term (powf):
    transcendental results may differ in the last bit, so nodes right on a
    sector boundary may differ."""
    nuc = [dict(i=12, j=12, k=12, Z=2, split="third", split_idx=s, split_axis=(0, 0, 1)) for s in range(3)]
    nuc += [dict(i=13, j=11, k=12, Z=2, split="tetra", split_idx=2)]
    mf, host = _pair(24, nuc, targets=[1, 1, 1, 2], USER_SHELL_MODE=True, USER_ALPHA_TF=0.3)
    with host:
        Ko, Uo, Wo, Po = O.init_fields(mf)
        _load(host, Ko, Uo, Wo, Po)
        host.do_steps(1)
        _, U, W, P = _fields(host, mf.NE)
        U1, W1, P1 = O.update(mf, Ko, Uo, Wo, Po)
        U1 = O.normalize(mf, U1, O.reduce_sums(mf, Ko, U1, W1, P1))
        assert ulp_diff(P, P1) == 0
        assert np.mean(W != W1) < 2e-3
        assert np.allclose(U, U1, rtol=1e-5, atol=1e-7)


def test_do_steps_schedule_matches_oracle():
    """65 steps: normalisation every 20 and at the end, 2-grid corrections at
    steps 25 and 50 (vcycle interval 25)."""
    mf, host = _pair(24, WATER, targets=[3, 1, 1], USER_VCYCLE_INTERVAL=25)
    with host:
        Ko, Uo, Wo, Po = O.init_fields(mf)
        _load(host, Ko, Uo, Wo, Po)
        host.do_steps(65)
        _, U, W, P = _fields(host, 3)
        U1, W1, P1, sums = O.do_steps(mf, Ko, Uo, Wo, Po, 65)
        assert np.allclose(host.sums, sums, rtol=1e-9)
        for a, b in ((U, U1), (W, W1), (P, P1)):
            assert np.allclose(a, b, rtol=1e-5, atol=1e-6), float(np.max(np.abs(a - b)))
        assert math.isfinite(host.E)
        assert host.E == pytest.approx(sums[3] + sums[4] + sums[5] + host.E_KK, rel=1e-12)


def test_forces_extract_recompute_k_match_oracle():
    mf, host = _pair(24, WATER, targets=[3, 1, 1])
    with host:
        host.do_steps(3)
        K, U, W, P = _fields(host, 3)
        fe = host.compute_forces()
        felec = O.forces_elec(mf, P)
        want = O.forces_total(mf, felec)
        assert np.allclose(np.array(fe), want, rtol=0, atol=1e-9)
        ex = host.extract()
        assert np.array_equal(ex, O.extract(mf, U, W, P))
        # fractional positions after a nuclear step, then the K rebuild
        host.nuclei[1]["i"] += 0.37
        host.nuclei[2]["j"] -= 0.21
        import ctypes as C
        pos = (C.c_double * 9)(*[x for n in host.nuclei for x in (n["i"], n["j"], n["k"])])
        host._check(host._lib.rq_mf_set_positions(host._h, pos))
        host._check(host._lib.rq_mf_recompute_k(host._h))
        atoms = [(f32(n["i"]), f32(n["j"]), f32(n["k"]), f32(n["Z"])) for n in host.nuclei]
        assert ulp_diff(host.field("K"), O.recompute_k(mf, atoms)) == 0


def test_frame_with_dynamics_moves_free_nuclei():
    nuc = [dict(i=12, j=12, k=12, Z=1), dict(i=16, j=12, k=12, Z=1, fixed=True)]
    cfg = {"USER_NN": 24, "USER_SCREEN": 8.0, "USER_NUCLEI": nuc, "USER_STEPS_PER_FRAME": 40,
           "USER_DYNAMICS": True, "USER_FORCE_SCALE": 50.0}
    with MolFast(cfg) as host:
        e0 = host.frame()
        before = (host.nuclei[0]["i"], host.nuclei[1]["i"])
        host.frame()
        assert math.isfinite(e0) and math.isfinite(host.E)
        assert host.nuclei[1]["i"] == before[1]  # fixed
        assert host.nuclei[0]["i"] != before[0] or host.nuclei[0]["j"] != 12
        assert host.E_KK == pytest.approx(O.e_kk(O.MolFast([O.Nucleus(**{k: n[k] for k in ("i", "j", "k", "Z")})
                                                            for n in host.nuclei], NN=24, screen=8.0)))


@pytest.mark.parametrize("NN,blocks", [(24, None), (70, None), (70, "180"), (130, None)])
def test_update_kernels_agree(monkeypatch, NN, blocks):
    """The three-plane LDS ring (default at three electrons, RQ_MF_LDS=2),
    the two-plane LDS march (1) and the per-node kernel (0) give the same
    fields over 30 steps (two normalisations), bit for bit: every kernel of
    the engine is deterministic (the reductions sum per-workgroup partials
    in a fixed order).  At NN = 70 the 69 interior
    nodes per axis leave ragged tiles: 5 of 64 k-columns and 1 of 4 j-rows
    in the two-plane march, 1 of 35 k-columns in the ring's 35 x 7 tiles
    (11 spare threads).  By default the ring splits each tile's planes into
    ranges of one or two planes; with RQ_MF_RING_GZ=5 and 180 target march
    workgroups the ring ranges hold 13 or 14 planes and the march blocks 14
    (the last 13).  At NN = 130 the ring tiles are 43 x 5 (three k-tiles
    over 129 nodes, 41 spare threads, 20 of them in a later wave than the
    node they repeat: their LDS writes go to the sink slot)."""
    if blocks is not None:
        monkeypatch.setenv("RQ_MF_BLOCKS", blocks)
        monkeypatch.setenv("RQ_MF_RING_GZ", "5")
    out = []
    c = NN // 2
    nuc = [dict(i=c, j=c, k=c, Z=3, rc=0.6), dict(i=c + 3, j=c + 2, k=c, Z=1), dict(i=c - 3, j=c + 2, k=c, Z=1)]
    for lds in ("2", "1", "0"):
        monkeypatch.setenv("RQ_MF_LDS", lds)
        mf, host = _pair(NN, nuc, targets=[3, 1, 1], USER_BETA_PAULI=0.2)
        with host:
            host.do_steps(30)
            out.append(_fields(host, 3))
    for other in out[1:]:
        for a, b in zip(out[0], other):
            assert ulp_diff(a, b) == 0


def test_seven_electrons_per_node_kernel_matches_oracle():
    """Above six electrons the per-node kernel runs: one step against the oracle."""
    nuc = [dict(i=8 + 2 * (a % 4), j=9 + 3 * (a // 4), k=10 + (a % 2), Z=1 + (a == 0), rc=0.4 * (a == 0))
           for a in range(7)]
    mf, host = _pair(20, nuc, targets=[2, 1, 1, 1, 1, 1, 1], screen=7.0)
    with host:
        Ko, Uo, Wo, Po = O.init_fields(mf)
        _load(host, Ko, Uo, Wo, Po)
        host.do_steps(1)
        _, U, W, P = _fields(host, 7)
        U1, W1, P1 = O.update(mf, Ko, Uo, Wo, Po)
        U1 = O.normalize(mf, U1, O.reduce_sums(mf, Ko, U1, W1, P1))
        assert ulp_diff(W, W1) == 0
        assert ulp_diff(P, P1) == 0
        assert ulp_diff(U, U1) <= 1
