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
