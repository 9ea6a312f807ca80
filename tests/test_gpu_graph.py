"""HIP-graph replay of step chunks gives the same trajectory as eager launches,
bit for bit: the kernels are the same and every reduction (norm, energy)
is an order-independent fixed-point sum, so two runs cannot differ."""
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
    assert ulp_diff(out[0][0], out[1][0]) == 0
    assert ulp_diff(out[0][1], out[1][1]) == 0
    assert np.array_equal(out[0][2], out[1][2])
    assert out[0][3] == out[1][3]


def test_repeated_runs_are_bit_identical():
    """Two eager runs of the same solve (normalisations, score rule, energy,
    HF forces) give identical fields and identical FP64 sums."""
    NN, screen = 40, 40 * 0.2
    rows = three_atoms(NN)
    out = []
    for _ in range(2):
        s = make_solver(rows, NN, screen, poisson="multigrid", mg_levels=3, bootstrap_poisson=5)
        s.initialize()
        s.step(0, 120)
        s._compute_energy()
        s._compute_forces()
        out.append((s.U.get(), s.E_domains.copy(), s.nuc_force.copy(), s.domain_norms()))
        s.close()
    assert ulp_diff(out[0][0], out[1][0]) == 0
    for a, b in zip(out[0][1:], out[1][1:]):
        assert np.array_equal(a, b)
