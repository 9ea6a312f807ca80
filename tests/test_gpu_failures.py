"""Failure detection through the C ABI, forced from the host.

molecule.js stops a run on a non-finite energy ("Numerical instability at
step N", molecule.js:3198-3201) and on a lost device (device.lost,
:2258-2261).  Here a NaN or Inf written into U or P must come back from
rq_step as RQ_EUNSTABLE at the next normalisation (RealQMSolver.run raises
NumericalInstability naming the step), and a HIP error injected with
rq_debug_fail_after must leave the handle refusing further work until
rq_destroy.  Non-convergence is forced by capping the relaxation's
iterations.  No kernel sees an out-of-range index: NaN only flows through
arithmetic."""
import ctypes as C

import numpy as np
import pytest

from realmolecule_amd import _lib as L
from tests.helpers import cube, make_solver, three_atoms

pytestmark = pytest.mark.gpu
NN = 24
SCREEN = NN * 0.3


def _solver(**cfg):
    s = make_solver(three_atoms(NN), NN, SCREEN, **cfg)
    s.initialize()
    return s


def _poison(s, field, value):
    S = NN + 1
    a = cube(getattr(s, field).get(), S).copy()
    a[NN // 2 - 4, NN // 2, NN // 2] = value
    getattr(s, field).set(a)


@pytest.mark.parametrize("graph", [False, True])
def test_nan_amplitude_stops_the_run(graph):
    s = _solver(graph=graph)
    s.step(0, 20)  # healthy steps pass the check
    _poison(s, "U", np.nan)
    print(f"NaN in U, graph={graph}", flush=True)
    with pytest.raises(L.NumericalInstability) as e:
        s.run(60, report_interval=20)
    assert "Numerical instability at step 20" in str(e.value)
    assert e.value.code == L.RQ_EUNSTABLE
    # the handle stays usable: a fresh trial clears the flag
    s.initialize()
    s.step(0, 40)
    s._compute_energy()
    assert np.isfinite(s.E)
    s.close()


def test_inf_potential_reaches_the_amplitudes_and_is_reported():
    s = _solver()
    _poison(s, "P", np.inf)
    print("Inf in P", flush=True)
    with pytest.raises(L.NumericalInstability) as e:
        s.step(0, 20, poisson_interval=1000)  # no Poisson step that could wash P out first
    assert "non-finite amplitude" in str(e.value)
    with pytest.raises(L.NumericalInstability):
        s._compute_energy()
    s.close()


def test_steps_without_a_normalisation_do_not_wait():
    """The check runs only when a normalisation falls in the call's range."""
    s = _solver()
    _poison(s, "U", np.nan)
    s.step(0, 19)  # steps 1..19: no normalisation, no error yet
    with pytest.raises(L.NumericalInstability):
        s.step(19, 1)
    s.close()


@pytest.mark.parametrize("nth", [1, 3])
def test_injected_hip_error_makes_the_handle_refuse_work(nth):
    s = _solver()
    lib, h = s._lib, s._h
    assert lib.rq_debug_fail_after(h, nth) == L.RQ_OK
    sched = L.RqSchedule(2, 20, 100)
    print(f"injected HIP error at check {nth}", flush=True)
    rc = lib.rq_step(h, 0, 40, C.byref(sched))
    assert rc == L.RQ_EHIP
    first = lib.rq_last_error(h).decode()
    assert "launch failure" in first.lower() or "hipErrorLaunchFailure" in first
    # every later call is refused with the first error in its message
    out = (C.c_double * 6)()
    assert lib.rq_energy(h, out, None) == L.RQ_EHIP
    msg = lib.rq_last_error(h).decode()
    assert "unusable" in msg and first in msg
    assert lib.rq_sync(h) == L.RQ_EHIP
    assert lib.rq_debug_fail_after(h, 0) == L.RQ_EHIP
    # destroy still drains the stream and frees the handle
    s.close()


def test_disarmed_injection_changes_nothing():
    a, b = _solver(), _solver()
    assert b._lib.rq_debug_fail_after(b._h, 0) == L.RQ_OK
    a.step(0, 40)
    b.step(0, 40)
    assert np.array_equal(a.U.get(), b.U.get())
    a.close()
    b.close()


def test_capped_relaxation_reports_non_convergence():
    from realmolecule_amd.relax import relax
    c = NN / 2
    rows = np.array([[c - 4.0, c, c, 1.0, 0.0], [c + 4.0, c, c, 1.0, 0.0]])
    out = relax(rows, NN, SCREEN, n_steps=20, bootstrap=5, max_iter=1, gtol=1e-12, checkpoint=16)
    assert not out["success"]
    assert "LIMIT" in out["message"].upper()
