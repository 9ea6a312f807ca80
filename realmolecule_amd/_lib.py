"""ctypes binding of ``librq.so`` (the C ABI declared in ``include/rq.h``).

The library is loaded from this package directory.  There is no fallback: if
it is missing or fails to load, importing the solver raises ``RuntimeError``
naming the build step.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librq.so")

RQ_OK = 0
RQ_EINVAL = -1
RQ_ENOMEM = -2
RQ_EHIP = -3
RQ_ESTATE = -4
RQ_ERANGE = -5
RQ_EUNSTABLE = -6

RQ_POISSON_JACOBI = 0
RQ_POISSON_MULTIGRID = 1
RQ_POISSON_TWOGRID = 2

RQ_BOUNDARY_SCORE = 0
RQ_BOUNDARY_LEVELSET = 1
RQ_BOUNDARY_OFF = 2

RQ_INIT_S1 = 0
RQ_INIT_S2 = 1
RQ_ENERGY_S1 = 0
RQ_ENERGY_S2 = 1
RQ_POTHER_TOTAL = 0
RQ_POTHER_DIRECT = 1
RQ_POTHER_FRAME = 2
RQ_POTHER_DIRECT_SWEEPS = 3

FIELDS = {"U": 0, "U2": 1, "K": 2, "P": 3, "P2": 4, "rho": 5, "label": 6, "bestU": 7, "W": 8, "Pother": 9}
FAMILIES = ["itp", "smooth", "smooth_c", "restrict", "restrict_c", "prolong", "prolong_c", "coarse_solve",
            "normalize", "boundary", "energy", "force", "smooth_itp", "bottom", "smooth_restrict", "post_itp"]


class RqAtom(C.Structure):
    _fields_ = [("i", C.c_float), ("j", C.c_float), ("k", C.c_float), ("Z", C.c_float), ("rc", C.c_float)]


class RqParams(C.Structure):
    _fields_ = [
        ("screen", C.c_double), ("dv", C.c_double), ("NN", C.c_int32), ("init_mode", C.c_int32),
        ("poisson_mode", C.c_int32), ("jacobi_sweeps", C.c_int32), ("smooth_a", C.c_float),
        ("smooth_c", C.c_float), ("mg_levels", C.c_int32), ("mg_nu1", C.c_int32), ("mg_nu2", C.c_int32),
        ("mg_a", C.c_float), ("mg_c", C.c_float), ("twogrid_interval", C.c_int32),
        ("twogrid_coarse_sweeps", C.c_int32), ("boundary_mode", C.c_int32), ("flip_ratio", C.c_float),
        ("flip_min_foreign", C.c_int32), ("ls_dt_w", C.c_float), ("ls_curv_reg", C.c_float),
        ("energy_mode", C.c_int32), ("rc_dirichlet", C.c_int32), ("pother_mode", C.c_int32),
        ("sic_interval", C.c_int32), ("reserved", C.c_int32 * 6),
    ]


class RqAtomS2(C.Structure):
    _fields_ = [("Z_nuc", C.c_float), ("init_zeff", C.c_float), ("init_rcut", C.c_float),
                ("split_type", C.c_int32), ("split_idx", C.c_int32), ("split_axis", C.c_float * 3),
                ("split_rot", C.c_float)]


RQ_SPLIT = {"sphere": 0, "hemi": 2, "third": 3}


class RqSchedule(C.Structure):
    _fields_ = [("poisson_interval", C.c_int32), ("norm_interval", C.c_int32), ("boundary_interval", C.c_int32)]


class RqMolecule(C.Structure):
    _fields_ = [("NN", C.c_int32), ("screen", C.c_double), ("n_atoms", C.c_int32), ("atoms", C.POINTER(RqAtom))]


class RqAdjointSpec(C.Structure):
    _fields_ = [("n_steps", C.c_int64), ("bootstrap", C.c_int32), ("checkpoint_interval", C.c_int32),
                ("schedule", RqSchedule), ("start", C.c_int32), ("reserved", C.c_int32 * 3)]


RQ_ADJ_START_TRIAL = 0
RQ_ADJ_START_SAVED = 1


RQ_MF_MAX_ELEC = 16
RQ_MF_SPLIT = {"sphere": 0, "hemi": 2, "third": 3, "tetra": 4, "hemi_third": 5}
RQ_MF_FIELDS = {"U": 0, "W": 1, "P": 2, "K": 3}


class RqMfNucleus(C.Structure):
    _fields_ = [("i", C.c_double), ("j", C.c_double), ("k", C.c_double), ("Z", C.c_double), ("rc", C.c_double),
                ("r_out", C.c_double), ("tilt_x", C.c_double), ("tilt_y", C.c_double), ("hs", C.c_double),
                ("z_init", C.c_double), ("target", C.c_double), ("axis", C.c_double * 3), ("split", C.c_int32),
                ("split_idx", C.c_int32), ("split_fix", C.c_int32), ("reserved", C.c_int32)]


class RqMfParams(C.Structure):
    _fields_ = [("NN", C.c_int32), ("shell_mode", C.c_int32), ("full_self", C.c_int32), ("vcycle", C.c_int32),
                ("vcycle_interval", C.c_int32), ("coarse_iters", C.c_int32), ("norm_interval", C.c_int32),
                ("reserved0", C.c_int32), ("screen", C.c_double), ("kinetic_coeff", C.c_double),
                ("kc_multi", C.c_double), ("beta_pauli", C.c_double), ("alpha_tf", C.c_double),
                ("field_x", C.c_double), ("reserved", C.c_int32 * 8)]


_P = C.c_void_p
_I = C.c_int32
_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)

# name -> (restype, argtypes)
PROTOTYPES = {
    "rq_abi_version": (C.c_int, []),
    "rq_last_error": (C.c_char_p, [_P]),
    "rq_default_params": (None, [C.POINTER(RqParams)]),
    "rq_default_schedule": (None, [C.POINTER(RqSchedule)]),
    "rq_dv_for_electrons": (C.c_double, [_I]),
    "rq_create": (C.c_int, [C.POINTER(RqParams), C.POINTER(RqAtom), _I, C.POINTER(_P)]),
    "rq_destroy": (C.c_int, [_P]),
    "rq_sync": (C.c_int, [_P]),
    "rq_debug_fail_after": (C.c_int, [_P, _I]),
    "rq_get_info": (C.c_int, [_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I), _D, _D]),
    "rq_set_atoms": (C.c_int, [_P, C.POINTER(RqAtom), _I]),
    "rq_get_atoms": (C.c_int, [_P, C.POINTER(RqAtom), _I]),
    "rq_set_lock_groups": (C.c_int, [_P, C.POINTER(C.c_int32), _I]),
    "rq_set_atoms_s2": (C.c_int, [_P, C.POINTER(RqAtomS2), _I]),
    "rq_init_trial": (C.c_int, [_P]),
    "rq_initialize": (C.c_int, [_P, _I]),
    "rq_poisson": (C.c_int, [_P, _I]),
    "rq_update": (C.c_int, [_P]),
    "rq_normalize": (C.c_int, [_P]),
    "rq_boundary": (C.c_int, [_P]),
    "rq_recompute_k": (C.c_int, [_P]),
    "rq_twogrid_correct": (C.c_int, [_P]),
    "rq_step": (C.c_int, [_P, C.c_int64, _I, C.POINTER(RqSchedule)]),
    "rq_set_graph": (C.c_int, [_P, _I]),
    "rq_set_fused": (C.c_int, [_P, _I]),
    "rq_set_tile": (C.c_int, [_P, _I, _I, _I, _I]),
    "rq_get_tile": (C.c_int, [_P, _I, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)]),
    "rq_energy": (C.c_int, [_P, _D, _D]),
    "rq_domain_norms": (C.c_int, [_P, _D]),
    "rq_forces_hf": (C.c_int, [_P, _I, _I, _D]),
    "rq_forces_hf_noalloc": (C.c_int, [_P, _I, _I, _D]),
    "rq_forces_gradp": (C.c_int, [_P, _I, _D]),
    "rq_read_field": (C.c_int, [_P, _I, _P, C.c_size_t]),
    "rq_write_field": (C.c_int, [_P, _I, _P, C.c_size_t]),
    "rq_swap_u": (C.c_int, [_P]),
    "rq_field_device_ptr": (C.c_int, [_P, _I, C.POINTER(_P)]),
    "rq_stream": (C.c_int, [_P, C.POINTER(_P)]),
    "rq_display_slice": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_float), C.c_size_t]),
    "rq_pother_frame": (C.c_int, [_P, C.c_int64]),
    "rq_solve_frozen": (C.c_int, [_P, C.POINTER(RqAdjointSpec), _D]),
    "rq_adjoint_gradient": (C.c_int, [_P, C.POINTER(RqAdjointSpec), _D, _D]),
    "rq_adjoint_save_start": (C.c_int, [_P]),
    "rq_batch_last_error": (C.c_char_p, [_P]),
    "rq_batch_create": (C.c_int, [C.POINTER(RqParams), C.POINTER(RqMolecule), _I, C.POINTER(_P)]),
    "rq_batch_destroy": (C.c_int, [_P]),
    "rq_batch_info": (C.c_int, [_P, C.POINTER(_I), C.POINTER(C.c_int64)]),
    "rq_batch_set_tiles": (C.c_int, [_P, _I, _I, C.POINTER(C.c_int32)]),
    "rq_batch_set_graph": (C.c_int, [_P, _I]),
    "rq_batch_initialize": (C.c_int, [_P, _I]),
    "rq_batch_step": (C.c_int, [_P, C.c_int64, _I, C.POINTER(RqSchedule)]),
    "rq_batch_sync": (C.c_int, [_P]),
    "rq_batch_energy": (C.c_int, [_P, _D]),
    "rq_batch_read_field": (C.c_int, [_P, _I, _I, _P, C.c_size_t]),
    "rq_batch_set_atoms": (C.c_int, [_P, _I, C.POINTER(RqAtom), _I]),
    "rq_batch_solve_frozen": (C.c_int, [_P, C.POINTER(RqAdjointSpec), _D]),
    "rq_batch_adjoint_gradient": (C.c_int, [_P, C.POINTER(RqAdjointSpec), _D, _D]),
    "rq_batch_timing_enable": (C.c_int, [_P, _I]),
    "rq_batch_timing_read": (C.c_int, [_P, _D, C.POINTER(C.c_int64), _I]),
    "rq_timing_enable": (C.c_int, [_P, _I]),
    "rq_timing_read": (C.c_int, [_P, _D, C.POINTER(C.c_int64), _I]),
    # include/rq_molfast.h
    "rq_mf_default_params": (None, [C.POINTER(RqMfParams)]),
    "rq_mf_last_error": (C.c_char_p, [_P]),
    "rq_mf_create": (C.c_int, [C.POINTER(RqMfParams), C.POINTER(RqMfNucleus), _I, C.POINTER(_P)]),
    "rq_mf_destroy": (C.c_int, [_P]),
    "rq_mf_sync": (C.c_int, [_P]),
    "rq_mf_init": (C.c_int, [_P]),
    "rq_mf_do_steps": (C.c_int, [_P, _I, _D]),
    "rq_mf_set_r_out": (C.c_int, [_P, _I]),
    "rq_mf_forces": (C.c_int, [_P, _F]),
    "rq_mf_set_positions": (C.c_int, [_P, _D]),
    "rq_mf_recompute_k": (C.c_int, [_P]),
    "rq_mf_extract": (C.c_int, [_P, _F, C.c_size_t]),
    "rq_mf_read_field": (C.c_int, [_P, _I, _I, _F, C.c_size_t]),
    "rq_mf_write_field": (C.c_int, [_P, _I, _I, _F, C.c_size_t]),
    "rq_mf_set_timing": (C.c_int, [_P, _I]),
    "rq_mf_update_time": (C.c_int, [_P, _D, C.POINTER(_I)]),
    "rq_mf_ring_tile": (None, [_I, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)]),
}

_lib = None


def load(path: str | None = None):
    """Load librq.so once and attach prototypes; raises RuntimeError if absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"librq.so not found at {p}: the HIP library is not built "
            "(run `python -c 'import __graft_entry__ as g; g.build()'` or `make -C realmolecule_amd/csrc`)")
    try:
        lib = C.CDLL(p)
    except OSError as e:  # pragma: no cover - depends on the ROCm runtime
        raise RuntimeError(f"failed to load {p}: {e}") from e
    for name, (res, args) in PROTOTYPES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib


class RqError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rq error {code}: {msg}")
        self.code = code


class NumericalInstability(RqError):
    """RQ_EUNSTABLE: a non-finite amplitude or energy (the reference page
    stops with "Numerical instability at step N", molecule.js:3198-3201)."""


def check(lib, h, code):
    if code != RQ_OK:
        msg = lib.rq_last_error(h)
        cls = NumericalInstability if code == RQ_EUNSTABLE else RqError
        raise cls(code, msg.decode() if msg else "")
    return code
