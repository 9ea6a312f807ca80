"""Batched solver: many molecules with uneven grids in one launch per pass.

The Python face of ``rq_batch_*`` (include/rq.h).  Replaces a geometry scan
that constructs and runs the reference's per-molecule solver once per
conformer (``cloud/solver.py`` ``RealQMSolver(config).run(...)``): all
conformers advance together, every pass of the step is one launch over
all their tiles (a workgroup per tile), and step chunks replay from HIP
graphs.
"""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

from . import _lib as L
from .solver import make_params, atoms_array


class BatchError(RuntimeError):
    pass


class Batch:
    def __init__(self, molecules, mg_levels=4, mg_nu=2, graph=True, dv=None, tiles=None):
        """molecules: list of (rows[n,5] grid units, NN, screen).  ``tiles``:
        optional (bx, by, {NN: ich}) launch shape, or 'tuned' for the
        autotuner's cached shapes of these grid sizes (autotune.batch_tiles)."""
        self._lib = L.load()
        self.NN = [int(m[1]) for m in molecules]
        self.screen = [float(m[2]) for m in molecules]
        self.rows = [np.asarray(m[0], np.float64) for m in molecules]
        self.n = len(molecules)
        cfg = {"poisson": "multigrid", "mg_levels": mg_levels, "mg_nu1": mg_nu, "mg_nu2": mg_nu}
        if dv is not None:
            cfg["dv"] = dv
        p = make_params(self.NN[0], self.screen[0], cfg)
        self._atoms = [atoms_array(r) for r in self.rows]
        mols = (L.RqMolecule * self.n)()
        for i in range(self.n):
            mols[i].NN = self.NN[i]
            mols[i].screen = self.screen[i]
            mols[i].n_atoms = len(self.rows[i])
            mols[i].atoms = C.cast(self._atoms[i], C.POINTER(L.RqAtom))
        h = C.c_void_p()
        rc = self._lib.rq_batch_create(C.byref(p), mols, self.n, C.byref(h))
        if rc:
            raise BatchError(f"rq_batch_create: {rc}: {self._lib.rq_batch_last_error(None).decode()}")
        self._h = h
        if graph:
            self._chk(self._lib.rq_batch_set_graph(h, 1))
        if tiles == "tuned":
            from .autotune import batch_tiles
            tiles = batch_tiles(self.NN)
        if tiles is not None:
            self.set_tiles(*tiles)

    def _chk(self, rc):
        if rc:
            raise BatchError(f"rq error {rc}: {self._lib.rq_batch_last_error(self._h).decode()}")

    def set_tiles(self, bx, by, ich_by_nn=None):
        arr = None
        if ich_by_nn:
            arr = (C.c_int32 * self.n)(*[int(ich_by_nn.get(nn, ich_by_nn.get(str(nn), 0))) for nn in self.NN])
        self._chk(self._lib.rq_batch_set_tiles(self._h, int(bx), int(by), arr))

    def interior_voxels(self):
        nm = C.c_int32()
        v = C.c_int64()
        self._chk(self._lib.rq_batch_info(self._h, C.byref(nm), C.byref(v)))
        return v.value

    def initialize(self, bootstrap=10):
        self._chk(self._lib.rq_batch_initialize(self._h, int(bootstrap)))

    def step(self, first, n, poisson_interval=1, norm_interval=20):
        sc = L.RqSchedule(int(poisson_interval), int(norm_interval), 1 << 30)
        self._chk(self._lib.rq_batch_step(self._h, int(first), int(n), C.byref(sc)))

    def sync(self):
        self._chk(self._lib.rq_batch_sync(self._h))

    def energies(self):
        """(n, 3) array of (T, V_eK, V_ee) per molecule, FP64."""
        out = (C.c_double * (3 * self.n))()
        self._chk(self._lib.rq_batch_energy(self._h, out))
        return np.frombuffer(out, np.float64).reshape(self.n, 3).copy()

    def read(self, m, field="U"):
        S = self.NN[m] + 1
        dt = np.int32 if field == "label" else np.float32
        out = np.empty(S ** 3, dt)
        self._chk(self._lib.rq_batch_read_field(self._h, int(m), L.FIELDS[field], out.ctypes.data, out.nbytes))
        return out

    def set_atoms(self, m, rows):
        """Positions (grid units) and rc of molecule m's atoms, for the frozen map."""
        rows = np.asarray(rows, np.float64)
        self._chk(self._lib.rq_batch_set_atoms(self._h, int(m), atoms_array(rows), len(rows)))

    def frozen_energies(self, spec):
        """(n, 3) (T, V_eK, V_ee) of the frozen-label map of every molecule."""
        e = (C.c_double * (3 * self.n))()
        self._chk(self._lib.rq_batch_solve_frozen(self._h, C.byref(spec), e))
        return np.frombuffer(e, np.float64).reshape(self.n, 3).copy()

    def adjoint_gradient(self, spec):
        """((n, 3) energies, [dE/dR (n_atoms_m, 3) per molecule]) of the
        frozen-label map, forward map and reverse sweep over all molecules at
        once (electronic part, Ha/bohr)."""
        na = sum(len(r) for r in self.rows)
        e = (C.c_double * (3 * self.n))()
        g = (C.c_double * (3 * na))()
        self._chk(self._lib.rq_batch_adjoint_gradient(self._h, C.byref(spec), e, g))
        flat = np.frombuffer(g, np.float64).reshape(na, 3).copy()
        out, a0 = [], 0
        for r in self.rows:
            out.append(flat[a0:a0 + len(r)])
            a0 += len(r)
        return np.frombuffer(e, np.float64).reshape(self.n, 3).copy(), out

    def timing(self, enable):
        self._chk(self._lib.rq_batch_timing_enable(self._h, int(enable)))

    def timing_read(self):
        nf = len(L.FAMILIES)
        ms = (C.c_double * nf)()
        cnt = (C.c_int64 * nf)()
        self._chk(self._lib.rq_batch_timing_read(self._h, ms, cnt, nf))
        return {L.FAMILIES[f]: (ms[f], cnt[f]) for f in range(nf) if cnt[f]}

    def close(self):
        if getattr(self, "_h", None) is not None:
            self._lib.rq_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
