"""HIP-graph replay of step chunks gives the same trajectory as eager launches."""
import numpy as np
import pytest

from tests.helpers import three_atoms, make_solver, ulp_diff

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("poisson", ["jacobi", "multigrid"])
def test_graph_replay_matches_eager(poisson):
    NN, screen = 32, 32 * 0.25
    rows = three_atoms(NN)
    out = []
    for graph in (False, True):
        s = make_solver(rows, NN, screen, poisson=poisson, mg_levels=3, graph=graph, bootstrap_poisson=5)
        s.initialize()
        pi = 1 if poisson == "multigrid" else 2
        s.step(0, 7, poisson_interval=pi)       # unaligned head runs eagerly
        s.step(7, 233, poisson_interval=pi)     # aligned 20-step chunks from graphs, boundary at 100 and 200
        s._compute_energy()
        out.append((s.U.get(), s.P.get(), s.label.get(), s.E))
        s.close()
    assert ulp_diff(out[0][0], out[1][0]) <= 1
    assert ulp_diff(out[0][1], out[1][1]) <= 1
    assert np.array_equal(out[0][2], out[1][2])
    assert out[0][3] == pytest.approx(out[1][3], rel=1e-12)
