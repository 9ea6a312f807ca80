"""The autotuner's launch shapes: a sweep on a small grid picks a shape, and
a solver or batch running the cached (or swept) shape evolves bit for bit
like one on the built-in shapes (each node's value does not depend on the
tiling; the reductions are order-independent fixed-point sums)."""
import numpy as np
import pytest

from realmolecule_amd import autotune
from realmolecule_amd.batch import Batch
from tests.helpers import make_solver, small_h2o

pytestmark = pytest.mark.gpu


def _run(NN, **cfg):
    rows, els, screen = small_h2o(NN, 0.1)
    s = make_solver(rows, NN, screen, poisson="multigrid", mg_levels=4, bootstrap_poisson=4, **cfg)
    s.initialize()
    return s


def test_cached_shape_is_applied_and_changes_nothing(tmp_path, monkeypatch):
    NN = 96
    # a cache entry unlike the built-in shape of 97 nodes per axis
    monkeypatch.setattr(autotune, "CACHE", str(tmp_path / "tiles.json"))
    autotune.save_cache({"gfx950:S97": {"bx": 128, "by": 2, "ich": 6, "ms_per_step": 0.0}})
    a, b = _run(NN), _run(NN, tiles="tuned")
    assert b.tiles_tuned and not a.tiles_tuned
    ent = autotune.load_cache()["gfx950:S97"]
    import ctypes as C
    bx, by, ich = C.c_int32(), C.c_int32(), C.c_int32()
    b._chk(b._lib.rq_get_tile(b._h, 0, C.byref(bx), C.byref(by), C.byref(ich)))
    assert (bx.value, by.value, ich.value) == (ent["bx"], ent["by"], ent["ich"])
    a.step(0, 40, poisson_interval=1)
    b.step(0, 40, poisson_interval=1)
    assert np.array_equal(a.U.get(), b.U.get())
    assert np.array_equal(a.P.get(), b.P.get())
    a._compute_energy()
    b._compute_energy()
    assert a.E == b.E
    a.close()
    b.close()


def test_sweep_picks_a_shape_and_keeps_results():
    NN = 48
    a, b = _run(NN), _run(NN)
    res = autotune.tune_solver(b, steps=2, shapes=[(64, 4), (32, 8)], ich=[1, 4], save=False)
    assert (res["bx"], res["by"]) in [(64, 4), (32, 8)] and res["ich"] in (1, 4)
    assert len(res["sweep"]) == 4 and all(t > 0 for t, *_ in res["sweep"])
    # the sweep advanced b by 4 x (2 + 2) steps far from the schedule; rerun
    # both from a fresh start with b's tuned shape kept
    a.initialize()
    b.initialize()
    a.step(0, 20, poisson_interval=1)
    b.step(0, 20, poisson_interval=1)
    assert np.array_equal(a.U.get(), b.U.get())
    a.close()
    b.close()


def test_batch_with_cached_shapes_matches_builtin(tmp_path, monkeypatch):
    monkeypatch.setattr(autotune, "CACHE", str(tmp_path / "tiles.json"))
    autotune.save_cache({"gfx950:S97": {"bx": 128, "by": 2, "ich": 6, "ms_per_step": 0.0}})
    rows, els, screen = small_h2o(96, 0.1)
    mols = [(rows, 96, screen), (rows, 96, screen)]
    x, y = Batch(mols, mg_levels=4, graph=False), Batch(mols, mg_levels=4, graph=False, tiles="tuned")
    for b in (x, y):
        b.initialize(bootstrap=4)
        b.step(0, 20)
    assert np.array_equal(x.read(0, "U"), y.read(0, "U"))
    assert np.array_equal(x.energies(), y.energies())
    x.close()
    y.close()
