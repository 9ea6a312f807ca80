"""Fused multigrid passes give bit-identical results to the unfused sequence
(which the V-cycle oracle tests pin)."""
import numpy as np
import pytest

from tests.helpers import three_atoms, make_solver, ulp_diff

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nu", [1, 2, 3])
def test_fused_vcycle_steps_bitexact(nu):
    NN, screen = 32, 32 * 0.25
    rows = three_atoms(NN)
    out = []
    for fused in (0, 1):
        s = make_solver(rows, NN, screen, poisson="multigrid", mg_levels=3, mg_nu1=nu, mg_nu2=nu,
                        bootstrap_poisson=3)
        s._chk(s._lib.rq_set_fused(s._h, fused))
        s.initialize()
        s.step(0, 45, poisson_interval=1)
        out.append((s.U.get(), s.P.get()))
        s.close()
    assert ulp_diff(out[0][0], out[1][0]) == 0
    assert ulp_diff(out[0][1], out[1][1]) == 0


@pytest.mark.parametrize("NN,levels,nu", [(96, 4, 2), (48, 3, 1), (64, 4, 3)])
def test_bottom_levels_bitexact(NN, levels, nu):
    """The single-workgroup launch of the last two levels (up to 25^3 nodes)
    against the same cycle with separate launches, fused (mode 1) and
    unfused (mode 0)."""
    screen = NN * 0.2
    rows = three_atoms(NN)
    out = []
    for mode in (0, 1, 3):
        s = make_solver(rows, NN, screen, poisson="multigrid", mg_levels=levels, mg_nu1=nu, mg_nu2=nu,
                        bootstrap_poisson=2)
        s._chk(s._lib.rq_set_fused(s._h, mode))
        s.initialize()
        s.step(0, 25, poisson_interval=1)
        out.append((s.U.get(), s.P.get()))
        s.close()
    for u, p in out[1:]:
        assert ulp_diff(out[0][0], u) == 0
        assert ulp_diff(out[0][1], p) == 0


def test_fused_large_grid_bitexact():
    """Levels above 63 coarse interior nodes per axis restrict with the
    i-march kernel (NN = 160: 79 coarse nodes), against the unfused passes."""
    NN, levels = 160, 4
    rows = three_atoms(NN)
    out = []
    for mode in (0, 1):
        s = make_solver(rows, NN, NN * 0.1, poisson="multigrid", mg_levels=levels, mg_nu1=2, mg_nu2=2,
                        bootstrap_poisson=2)
        s._chk(s._lib.rq_set_fused(s._h, mode))
        s.initialize()
        s.step(0, 12, poisson_interval=1)
        out.append((s.U.get(), s.P.get()))
        s.close()
    assert ulp_diff(out[0][0], out[1][0]) == 0
    assert ulp_diff(out[0][1], out[1][1]) == 0


@pytest.mark.parametrize("NN,levels,nu2", [(96, 4, 2), (64, 4, 1), (160, 4, 3), (48, 3, 2)])
def test_level_legs_bitexact(NN, levels, nu2):
    """One down-leg and one up-leg launch per coarse level (two pre-sweeps,
    nu2 post-sweeps) against the separate fused passes (mode 2) and the
    unfused sequence (mode 0)."""
    rows = three_atoms(NN)
    out = []
    for mode in (0, 2, 1):
        s = make_solver(rows, NN, NN * 0.15, poisson="multigrid", mg_levels=levels, mg_nu1=2, mg_nu2=nu2,
                        bootstrap_poisson=2)
        s._chk(s._lib.rq_set_fused(s._h, mode))
        s.initialize()
        s.step(0, 15, poisson_interval=1)
        out.append((s.U.get(), s.P.get()))
        s.close()
    for u, p in out[1:]:
        assert ulp_diff(out[0][0], u) == 0
        assert ulp_diff(out[0][1], p) == 0


@pytest.mark.parametrize("NN,sweeps,min_s", [(48, 2, 0), (100, 5, 0), (130, 4, 0), (320, 2, None)])
def test_sweep_pairs_jacobi_bitexact(NN, sweeps, min_s, monkeypatch):
    """Jacobi sweep pairs run as one two-sweep march (fused) and equal the
    one-sweep launches bit for bit, on grids that are no multiple of the
    march's 64 x 8 column tile, with an odd count leaving one single sweep.
    Small grids lower the size gate of the march (RQ_SMOOTH2_MIN_S=0);
    the 321^3 grid takes it by default."""
    if min_s is not None:
        monkeypatch.setenv("RQ_SMOOTH2_MIN_S", str(min_s))
    rows = three_atoms(NN)
    out = []
    for mode in (0, 1):
        s = make_solver(rows, NN, NN * 0.2, poisson="jacobi", jacobi_sweeps=sweeps)
        s._chk(s._lib.rq_set_fused(s._h, mode))
        s.initialize()
        s.step(0, 30, poisson_interval=1)
        out.append((s.U.get(), s.P.get()))
        s.close()
    assert ulp_diff(out[0][0], out[1][0]) == 0
    assert ulp_diff(out[0][1], out[1][1]) == 0


@pytest.mark.parametrize("NN,levels,nu1,nu2", [(136, 4, 3, 4), (192, 4, 2, 2)])
def test_sweep_pairs_multigrid_bitexact(NN, levels, nu1, nu2, monkeypatch):
    """Fine and coarse sweep pairs of the V-cycle as two-sweep marches
    (pre: nu1 sweeps, post: nu2 - 1 plain sweeps before the one fused with
    the ITP update; coarse post pairs on levels above the per-level legs)
    against the unfused sequence, with the march's size gate lowered."""
    monkeypatch.setenv("RQ_SMOOTH2_MIN_S", "0")
    rows = three_atoms(NN)
    out = []
    for mode in (0, 1):
        s = make_solver(rows, NN, NN * 0.12, poisson="multigrid", mg_levels=levels, mg_nu1=nu1, mg_nu2=nu2,
                        bootstrap_poisson=2)
        s._chk(s._lib.rq_set_fused(s._h, mode))
        s.initialize()
        s.step(0, 8, poisson_interval=1)
        out.append((s.U.get(), s.P.get()))
        s.close()
    assert ulp_diff(out[0][0], out[1][0]) == 0
    assert ulp_diff(out[0][1], out[1][1]) == 0


@pytest.mark.parametrize("NN,levels,blocks", [(80, 3, 0), (136, 4, 0), (136, 4, 1), (192, 4, 0)])
def test_presmooth_restrict_march_bitexact(NN, levels, blocks, monkeypatch):
    """The fine pre-smoothing pair, residual and restriction as one march
    (k_smooth2_resres_march; size gate lowered; ``blocks`` = 1 gives each
    workgroup every coarse plane) against the unfused sequence and against
    the pair + restriction march as two launches."""
    import ctypes as C
    from realmolecule_amd import _lib as L
    monkeypatch.setenv("RQ_SMOOTH2_MIN_S", "0")
    if blocks:
        monkeypatch.setenv("RQ_SMOOTH2_RESRES_BLOCKS", str(blocks))
    rows = three_atoms(NN)
    out = []
    for mode, s2rr in ((0, "1"), (1, "0"), (1, "1")):
        monkeypatch.setenv("RQ_SMOOTH2_RESRES", s2rr)
        s = make_solver(rows, NN, NN * 0.12, poisson="multigrid", mg_levels=levels, bootstrap_poisson=2)
        s._chk(s._lib.rq_set_fused(s._h, mode))
        s.initialize()
        s._chk(s._lib.rq_timing_enable(s._h, 1))
        s.step(0, 8, poisson_interval=1)
        nf = len(L.FAMILIES)
        ms, cnt = (C.c_double * nf)(), (C.c_int64 * nf)()
        s._chk(s._lib.rq_timing_read(s._h, ms, cnt, nf))
        s._chk(s._lib.rq_timing_enable(s._h, 0))
        fused_launches = cnt[L.FAMILIES.index("smooth_restrict")]
        out.append((s.U.get(), s.P.get(), fused_launches))
        s.close()
    assert out[0][2] == 0 and out[1][2] == 0 and out[2][2] == 8
    for k in (1, 2):
        assert ulp_diff(out[0][0], out[k][0]) == 0
        assert ulp_diff(out[0][1], out[k][1]) == 0


@pytest.mark.parametrize("NN,levels,blocks,rc,graph", [(80, 3, 0, False, False), (136, 4, 0, True, False),
                                                       (136, 4, 1, False, True), (192, 4, 0, True, True)])
def test_post_march_bitexact(NN, levels, blocks, rc, graph, monkeypatch):
    """The fine correction, post-smoothing pair and ITP update as one march
    (k_post_march; size gate lowered; ``blocks`` = 1 gives each workgroup
    every plane; with the rc rule's label flags; with graph replay) against
    the unfused sequence and against the separate fused launches."""
    import ctypes as C
    from realmolecule_amd import _lib as L
    monkeypatch.setenv("RQ_SMOOTH2_MIN_S", "0")
    if blocks:
        monkeypatch.setenv("RQ_POST_BLOCKS", str(blocks))
    rows = three_atoms(NN)
    rows[1, 4] = 0.6
    out = []
    for mode, post in ((0, "1"), (1, "0"), (1, "1")):
        monkeypatch.setenv("RQ_POST_FUSE", post)
        s = make_solver(rows, NN, NN * 0.12, poisson="multigrid", mg_levels=levels, bootstrap_poisson=2,
                        rc_dirichlet=rc, graph=graph)
        s._chk(s._lib.rq_set_fused(s._h, mode))
        s.initialize()
        s._chk(s._lib.rq_timing_enable(s._h, 1))
        s.step(0, 8, poisson_interval=1)
        nf = len(L.FAMILIES)
        ms, cnt = (C.c_double * nf)(), (C.c_int64 * nf)()
        s._chk(s._lib.rq_timing_read(s._h, ms, cnt, nf))
        s._chk(s._lib.rq_timing_enable(s._h, 0))
        post_launches = cnt[L.FAMILIES.index("post_itp")]
        if graph:
            s.step(8, 40, poisson_interval=1)  # replayed chunks
        out.append((s.U.get(), s.P.get(), post_launches))
        s.close()
    assert out[0][2] == 0 and out[1][2] == 0 and out[2][2] == 8
    for k in (1, 2):
        assert ulp_diff(out[0][0], out[k][0]) == 0
        assert ulp_diff(out[0][1], out[k][1]) == 0
