"""The C + OpenMP step (oracle/cstep.c, the all-core CPU baseline) against
the NumPy oracle: the same 25 steps of the multigrid schedule (V-cycle every
step, ITP, normalise at step 20) on the same inputs, in FP32."""
import numpy as np
import pytest

from oracle import cstep, mg, s1
from tests.helpers import three_atoms, ulp_diff


@pytest.mark.parametrize("NN,levels", [(24, 3), (32, 4)])
def test_cstep_matches_numpy_oracle(NN, levels):
    screen = NN * 0.25
    rows = three_atoms(NN)
    g = s1.Grid(NN, screen, 0.12, np.float32)
    U, lab, K, _ = s1.init_trial(rows, g)
    Z = rows[:, 3]
    U, _ = s1.normalize(U, lab, Z, g)
    P = np.zeros_like(U)
    Uc, Pc = U.copy(), P.copy()
    cstep.run_steps(Uc, Pc, K, lab.astype(np.int32), Z, g, levels, 0, 25)
    a, c = np.float32(1.0 / 7.0), np.float32(7.0)
    for step in range(1, 26):
        rhs = s1.rho_rhs(U, g.two_pi, g.h2, np.float32)
        P = mg.vcycle(P, rhs, levels, 2, 2, float(a), float(c), np.float32)
        U = s1.itp_update(U, K, P, lab, g)
        if step % 20 == 0:
            U, _ = s1.normalize(U, lab, Z, g)
    dU, dP = ulp_diff(Uc, U), ulp_diff(Pc, P)
    print(f"NN={NN}: U max ulp {dU}, P max ulp {dP}, threads {cstep.threads()}")
    assert np.allclose(Uc, U, rtol=2e-6, atol=1e-9)
    assert np.allclose(Pc, P, rtol=2e-6, atol=1e-9)


def test_cstep_rejects_bad_levels():
    NN = 24
    g = s1.Grid(NN, 6.0, 0.12, np.float32)
    S = NN + 1
    f = np.zeros((S, S, S), np.float32)
    with pytest.raises(ValueError):
        cstep.run_steps(f.copy(), f.copy(), f, np.zeros((S, S, S), np.int32), [1.0], g, 5, 0, 1)
