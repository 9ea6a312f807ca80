"""C ABI surface (CPU only): the library loads and exports every function
declared in include/rq.h; no compute call is made."""
import ctypes as C
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not fn.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", fn)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names += re.findall(r"^\s*(?:int|void|double|const char\*)\s+\**(rq\w+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_prototypes_cover_header():
    from realmolecule_amd import _lib
    names = declared_functions()
    assert set(names) <= set(_lib.PROTOTYPES), sorted(set(names) - set(_lib.PROTOTYPES))


def test_default_params_and_struct_layout(lib):
    from realmolecule_amd import _lib
    p = _lib.RqParams()
    lib.rq_default_params(C.byref(p))
    assert (p.NN, p.jacobi_sweeps, p.smooth_a, p.smooth_c) == (96, 2, 0.0, 6.0)
    assert abs(p.flip_ratio - 1.1) < 1e-7 and p.flip_min_foreign == 2
    assert C.sizeof(_lib.RqParams) == 2 * 8 + 20 * 4 + 8 * 4
    s = _lib.RqSchedule()
    lib.rq_default_schedule(C.byref(s))
    assert (s.poisson_interval, s.norm_interval, s.boundary_interval) == (2, 20, 100)
    assert lib.rq_abi_version() == 1


def test_dv_table_matches_reference(lib):
    # cloud/solver.py:341-351
    for n, dv in [(1, 0.12), (30, 0.12), (31, 0.03), (100, 0.03), (101, 0.02), (200, 0.02), (201, 0.01),
                  (500, 0.01), (501, 0.005)]:
        assert lib.rq_dv_for_electrons(n) == dv


def test_create_rejects_bad_params_without_gpu(lib):
    # argument validation runs before any device call
    from realmolecule_amd import _lib
    p = _lib.RqParams()
    lib.rq_default_params(C.byref(p))
    p.NN = 7
    atoms = (_lib.RqAtom * 1)()
    h = C.c_void_p()
    assert lib.rq_create(C.byref(p), atoms, 1, C.byref(h)) == _lib.RQ_EINVAL
    assert b"NN=7" in lib.rq_last_error(None)
    p.NN = 96
    p.poisson_mode = _lib.RQ_POISSON_MULTIGRID
    p.mg_levels = 2  # coarsest interior 47 > 25
    assert lib.rq_create(C.byref(p), atoms, 1, C.byref(h)) == _lib.RQ_ERANGE
    p.mg_levels = 4
    p.reserved[0] = 1
    assert lib.rq_create(C.byref(p), atoms, 1, C.byref(h)) == _lib.RQ_EINVAL
    # an oversize grid is refused before any allocation (the buffer-size
    # guard of molecule.js:2246-2250)
    p.reserved[0] = 0
    p.poisson_mode = _lib.RQ_POISSON_JACOBI
    p.NN = 2048
    assert lib.rq_create(C.byref(p), atoms, 1, C.byref(h)) == _lib.RQ_EINVAL
    assert b"NN=2048" in lib.rq_last_error(None)


def test_null_handle_is_refused_without_gpu(lib):
    """Every entry point that takes a handle refuses a null one before any
    device call (rq_destroy accepts it as a no-op)."""
    from realmolecule_amd import _lib
    sched = _lib.RqSchedule(2, 20, 100)
    out = (C.c_double * 6)()
    assert lib.rq_step(None, 0, 1, C.byref(sched)) == _lib.RQ_EINVAL
    assert lib.rq_init_trial(None) == _lib.RQ_EINVAL
    assert lib.rq_sync(None) == _lib.RQ_EINVAL
    assert lib.rq_energy(None, out, None) == _lib.RQ_EINVAL
    assert lib.rq_debug_fail_after(None, 1) == _lib.RQ_EINVAL
    assert b"null handle" in lib.rq_last_error(None)
    assert lib.rq_destroy(None) == _lib.RQ_OK
