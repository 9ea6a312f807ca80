"""The allocation-free HF force entry point (rq_forces_hf_noalloc) against
rq_forces_hf through the C ABI: identical results (the same kernel and
order-independent sums) on two calls in a row, the second on a smaller
molecule and grid, so the shared scratch array must be cleared between
calls; counts above its limit are rejected."""
import ctypes as C

import numpy as np
import pytest

from realmolecule_amd import _lib as L
from tests.helpers import three_atoms, make_solver

pytestmark = pytest.mark.gpu


def _both(s, first, count):
    lib = s._lib
    ref = (C.c_double * (3 * count))()
    got = (C.c_double * (3 * count))()
    s._chk(lib.rq_forces_hf(s._h, first, count, ref))
    s._chk(lib.rq_forces_hf_noalloc(s._h, first, count, got))
    return np.frombuffer(ref, np.float64).reshape(count, 3), np.frombuffer(got, np.float64).reshape(count, 3)


def test_noalloc_forces_match_allocating_entry_point():
    NN = 40
    rows = three_atoms(NN)
    a = make_solver(rows, NN, NN * 0.2, bootstrap_poisson=10)
    a.initialize()
    a.step(0, 40)
    ref, got = _both(a, 0, 3)
    print(f"40^3, 3 atoms: max |F| {np.abs(ref).max():.4f}", flush=True)
    assert np.abs(ref).max() > 1e-3
    assert np.array_equal(ref, got)
    # second call: a smaller molecule on a smaller grid, fewer atoms
    NN2 = 24
    rows2 = three_atoms(NN2)[:2]
    b = make_solver(rows2, NN2, NN2 * 0.3, bootstrap_poisson=10)
    b.initialize()
    b.step(0, 20)
    ref2, got2 = _both(b, 0, 2)
    assert np.array_equal(ref2, got2)
    ref3, got3 = _both(b, 1, 1)
    assert np.array_equal(ref3, got3)
    assert np.array_equal(ref3, ref2[1:])
    # and the first handle again after the smaller call
    ref4, got4 = _both(a, 0, 3)
    assert np.array_equal(got4, ref)
    a.close()
    b.close()


def test_noalloc_rejects_counts_beyond_its_scratch():
    NN = 16
    rng = np.random.default_rng(0)
    n = 1100
    rows = np.zeros((n, 5))
    rows[:, :3] = rng.uniform(2, NN - 2, (n, 3))
    rows[:, 3] = 1.0
    s = make_solver(rows, NN, NN * 0.3, bootstrap_poisson=0)
    out = (C.c_double * (3 * n))()
    code = s._lib.rq_forces_hf_noalloc(s._h, 0, n, out)
    assert code == L.RQ_ERANGE
    assert b"1024" in s._lib.rq_last_error(s._h)
    s.close()
