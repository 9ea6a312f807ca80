"""``RealQMSolver``: drop-in for ``cloud/solver.py``'s class of the same name.

Same constructor dict, same public methods (``initialize``, ``run``,
``set_fold_atoms``), the same private step methods callers use
(``_normalize``, ``_poisson_step``, ``_compute_energy``, ``_compute_forces``,
``_move_nuclei``) and the same attributes (``cloud/solver.py:321-400``).  The
device arrays ``U, U2, P, P2, K_buf, rho, label, bestU`` are
:class:`DeviceField` objects whose ``.get()`` returns a NumPy copy, like the
CuPy arrays callers read with ``.get()`` (``cloud/run_h2o_debug.py:20-57``);
``solver.U, solver.U2 = solver.U2, solver.U`` swaps the ping-pong pair as in
``cloud/benchmark.py:104``.

All grid work runs in ``librq.so`` (HIP, gfx950).  Host-side logic kept in
Python exactly where the reference keeps it on the host: nucleus-nucleus
energy and forces and the damped nuclear integrator (FP64 here; the
reference accumulates them in FP32 NumPy scalars).

Extra constructor keys (all optional; defaults reproduce cloud/solver.py):
``poisson`` ('jacobi' | 'multigrid' | 'twogrid'), ``mg_levels``, ``mg_nu1``,
``mg_nu2``, ``mg_omega``, ``jacobi_sweeps``, ``smooth_a``, ``smooth_c``,
``boundary`` ('score' | 'levelset' | 'off'), ``init`` ('s1' | 's2'),
``energy`` ('s1' | 's2'), ``rc_dirichlet`` (bool), ``dv``, ``graph`` (bool: replay
step chunks from HIP graphs), ``tiles`` ('tuned': the autotuner's cached launch
shape for the grid size), ``bootstrap_poisson`` (Poisson steps in initialize, 100),
``pother`` ('total' | 'direct' | 'frame' | 'direct_sweeps': the potential the
ITP update and the energy see) and ``sic_interval`` (self-interaction
correction every n frames with 'frame').  Atom dicts may carry the
molecule.js per-atom keys ``Z_nuc``, ``init_zeff``, ``init_rcut``,
``split`` ('sphere' | 'hemi' | 'third'), ``split_idx``, ``split_axis``,
``split_rot`` (used with ``init: 's2'``).
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib as L


class DeviceField:
    """Handle to one device field of a solver (a stand-in for a CuPy array)."""

    def __init__(self, solver: "RealQMSolver", name: str):
        self._s = solver
        self.name = name

    @property
    def shape(self):
        return (self._s.S3,)

    @property
    def dtype(self):
        return np.dtype(np.int32 if self.name == "label" else np.float32)

    @property
    def size(self):
        return self._s.S3

    def get(self) -> np.ndarray:
        """Copy to the host (flat, length S^3, like the reference's flat CuPy arrays)."""
        out = np.empty(self._s.S3, dtype=self.dtype)
        self._s._read(self.name, out)
        return out

    def set(self, values) -> None:
        arr = np.ascontiguousarray(np.asarray(values, dtype=self.dtype).reshape(-1))
        if arr.size != self._s.S3:
            raise ValueError(f"{self.name}: need {self._s.S3} values, got {arr.size}")
        self._s._write(self.name, arr)

    def device_ptr(self) -> int:
        p = C.c_void_p()
        L.check(self._s._lib, self._s._h, self._s._lib.rq_field_device_ptr(self._s._h, L.FIELDS[self.name], C.byref(p)))
        return p.value

    def __array__(self, dtype=None, copy=None):
        a = self.get()
        return a if dtype is None else a.astype(dtype)

    def __repr__(self):
        return f"DeviceField({self.name}, S^3={self._s.S3})"


_POISSON = {"jacobi": L.RQ_POISSON_JACOBI, "multigrid": L.RQ_POISSON_MULTIGRID, "twogrid": L.RQ_POISSON_TWOGRID}
_BOUNDARY = {"score": L.RQ_BOUNDARY_SCORE, "levelset": L.RQ_BOUNDARY_LEVELSET, "off": L.RQ_BOUNDARY_OFF}
_INIT = {"s1": L.RQ_INIT_S1, "s2": L.RQ_INIT_S2}
_ENERGY = {"s1": L.RQ_ENERGY_S1, "s2": L.RQ_ENERGY_S2}


def make_params(NN, screen, config=None) -> L.RqParams:
    """rq_params for a config dict (defaults = cloud/solver.py semantics)."""
    config = config or {}
    lib = L.load()
    p = L.RqParams()
    lib.rq_default_params(C.byref(p))
    p.NN = int(NN)
    p.screen = float(screen)
    if config.get("dv") is not None:
        p.dv = float(config["dv"])
    p.poisson_mode = _POISSON[config.get("poisson", "jacobi")]
    p.jacobi_sweeps = int(config.get("jacobi_sweeps", 2))
    if "smooth_a" in config:
        p.smooth_a = float(config["smooth_a"])
    if "smooth_c" in config:
        p.smooth_c = float(config["smooth_c"])
    p.mg_levels = int(config.get("mg_levels", 4))
    p.mg_nu1 = int(config.get("mg_nu1", 2))
    p.mg_nu2 = int(config.get("mg_nu2", 2))
    if "mg_omega" in config:
        w = float(config["mg_omega"])
        # x' = (1-w) x + w (sum6 + g)/6  ==  a x + (sum6 + g)/c with a = 1-w, c = 6/w
        p.mg_a = 1.0 - w
        p.mg_c = 6.0 / w
    p.twogrid_interval = int(config.get("twogrid_interval", 50))
    p.twogrid_coarse_sweeps = int(config.get("twogrid_coarse_sweeps", 10))
    p.boundary_mode = _BOUNDARY[config.get("boundary", "score")]
    p.init_mode = _INIT[config.get("init", "s1")]
    p.energy_mode = _ENERGY[config.get("energy", "s1")]
    p.rc_dirichlet = int(bool(config.get("rc_dirichlet", False)))
    # 'total': cloud/solver.py's potential of all electrons; 'direct': the
    # molecule.js direct Pother (each domain sees the other domains' charge)
    p.pother_mode = {"total": L.RQ_POTHER_TOTAL, "direct": L.RQ_POTHER_DIRECT, "frame": L.RQ_POTHER_FRAME,
                     "direct_sweeps": L.RQ_POTHER_DIRECT_SWEEPS}[config.get("pother", "total")]
    p.sic_interval = int(config.get("sic_interval", 0))
    if "flip_ratio" in config:
        p.flip_ratio = float(config["flip_ratio"])
    if "ls_dt_w" in config:
        p.ls_dt_w = float(config["ls_dt_w"])
    if "ls_curv_reg" in config:
        p.ls_curv_reg = float(config["ls_curv_reg"])
    return p


S2_KEYS = ("Z_nuc", "init_zeff", "init_rcut", "split", "split_idx", "split_axis", "split_rot")


def atoms_s2_array(atoms):
    """rq_atom_s2 rows from config atom dicts (molecule.js USER_Z_NUC,
    USER_INIT_ZEFF, USER_INIT_RCUT, USER_SPLIT_*); missing keys take the
    library defaults (0: Z_nuc = Z, init_zeff = Z, init_rcut = 1e6)."""
    arr = (L.RqAtomS2 * len(atoms))()
    for n, a in enumerate(atoms):
        arr[n].Z_nuc = float(a.get("Z_nuc", 0.0))
        arr[n].init_zeff = float(a.get("init_zeff", 0.0))
        arr[n].init_rcut = float(a.get("init_rcut", 0.0))
        st = a.get("split", 0)
        arr[n].split_type = L.RQ_SPLIT[st] if isinstance(st, str) else int(st)
        arr[n].split_idx = int(a.get("split_idx", 0))
        ax = a.get("split_axis", (0.0, 0.0, 1.0))
        for c in range(3):
            arr[n].split_axis[c] = float(ax[c])
        arr[n].split_rot = float(a.get("split_rot", 0.0))
    return arr


def atoms_array(atoms):
    arr = (L.RqAtom * len(atoms))()
    for n, a in enumerate(atoms):
        arr[n].i, arr[n].j, arr[n].k, arr[n].Z, arr[n].rc = [float(x) for x in a]
    return arr


class RealQMSolver:
    """MI355X solver with the constructor and methods of cloud/solver.py:318-638."""

    def __init__(self, config):
        self._lib = L.load()
        self._h = None
        self.NN = config['gridN']
        self.S = self.NN + 1
        self.S3 = self.S ** 3
        self.screen = config['screen']
        self.h = self.screen / self.NN
        self.h2 = self.h * self.h
        self.h3 = self.h ** 3
        self.N2 = self.NN // 2

        atoms = config['atoms']
        self.nAtoms = len(atoms)
        self.Z_eff = np.array([a['Z'] for a in atoms], dtype=np.float32)
        self.atom_pos = np.array([[a['i'], a['j'], a.get('k', self.N2)] for a in atoms], dtype=np.float32)
        self.atom_rc = np.array([a.get('rc', 0) for a in atoms], dtype=np.float32)
        self.atom_el = [a.get('el', '') for a in atoms]
        self.active = self.Z_eff > 0
        self.nElec = int(self.active.sum())

        self.dv = float(config['dv']) if config.get('dv') else float(self._lib.rq_dv_for_electrons(self.nElec))
        self.half_dv = 0.5 * self.dv
        self.dtv = self.dv * self.h2

        self.force_scale = config.get('forceScale', 1.0)
        self.damping = config.get('damping', 0.98)
        self.dt_nuc = config.get('dt_nuc', 0.8 if self.nElec <= 5 else 0.2 if self.nElec > 200 else 0.8)
        self.max_vel = 0.3 if self.nElec <= 5 else 0.03 if self.nElec > 200 else 0.1
        self.dynamics_enabled = config.get('dynamics', False)
        self._nProtein = config.get('nProtein', self.nAtoms)
        self._config_force_interval = config.get('force_interval', 500)

        self.nuc_vel = np.zeros((self.nAtoms, 3), dtype=np.float32)
        self.nuc_force = np.zeros((self.nAtoms, 3), dtype=np.float32)

        self._params = make_params(self.NN, self.screen, config)
        self._params.dv = self.dv
        h = C.c_void_p()
        code = self._lib.rq_create(C.byref(self._params), self._atoms_c(), self.nAtoms, C.byref(h))
        L.check(self._lib, None, code)
        self._h = h
        if any(k in a for a in atoms for k in S2_KEYS):
            self._chk(self._lib.rq_set_atoms_s2(self._h, atoms_s2_array(atoms), self.nAtoms))
        self._sched = L.RqSchedule()
        self._lib.rq_default_schedule(C.byref(self._sched))
        if config.get('graph'):
            self._chk(self._lib.rq_set_graph(self._h, 1))
        # 'tuned': the autotuner's cached fine-level launch shape for this grid
        # size, when one exists (results do not depend on the shape)
        self.tiles_tuned = False
        if config.get('tiles') == 'tuned':
            from .autotune import apply_cached
            self.tiles_tuned = apply_cached(self)

        self.E_T = 0.0
        self.E_eK = 0.0
        self.E_ee = 0.0
        self.E_KK = 0.0
        self.E = 0.0
        self.dipole = np.zeros(3)
        self.E_domains = np.zeros((self.nAtoms, 3))

        self._atom_ref = {'H': -0.5, 'O': -2.04, 'C': -5.43, 'N': -3.64}
        self.E_atoms_sum = sum(self._atom_ref.get(a.get('el', ''), 0) for a in atoms if a['Z'] > 0)
        self._initialized = False
        self._step = 0
        self._bootstrap = int(config.get('bootstrap_poisson', 100))

        self._fields = {n: DeviceField(self, n) for n in L.FIELDS}

    # ---- device fields (cloud/solver.py:365-372 attribute names) ----
    def _field(self, n):
        return self._fields[n]

    @property
    def U(self):
        return self._fields["U"]

    @U.setter
    def U(self, v):
        self._assign("U", v)

    @property
    def U2(self):
        return self._fields["U2"]

    @U2.setter
    def U2(self, v):
        self._assign("U2", v)

    @property
    def P(self):
        return self._fields["P"]

    @P.setter
    def P(self, v):
        self._assign("P", v)

    @property
    def P2(self):
        return self._fields["P2"]

    @P2.setter
    def P2(self, v):
        self._assign("P2", v)

    @property
    def Pother(self):
        """Frame potential of pother 'frame' (molecule.js PotherBuf)."""
        return self._fields["Pother"]

    @property
    def K_buf(self):
        return self._fields["K"]

    @K_buf.setter
    def K_buf(self, v):
        self._assign("K", v)

    @property
    def rho(self):
        return self._fields["rho"]

    @property
    def label(self):
        return self._fields["label"]

    @label.setter
    def label(self, v):
        self._assign("label", v)

    @property
    def bestU(self):
        return self._fields["bestU"]

    @property
    def W(self):
        return self._fields["W"]

    def _assign(self, name, v):
        if isinstance(v, DeviceField) and v._s is self:
            if {name, v.name} == {"U", "U2"}:
                # tuple-swap ``s.U, s.U2 = s.U2, s.U`` arrives as two assignments;
                # swap once on the first of them.
                if getattr(self, "_swap_pending", None) == (v.name, name):
                    self._swap_pending = None
                    return
                L.check(self._lib, self._h, self._lib.rq_swap_u(self._h))
                self._swap_pending = (name, v.name)
                return
            if v.name == name:
                return
            v = v.get()
        self._fields[name].set(v)

    def _read(self, name, out):
        L.check(self._lib, self._h, self._lib.rq_read_field(self._h, L.FIELDS[name], out.ctypes.data, out.nbytes))

    def _write(self, name, arr):
        L.check(self._lib, self._h, self._lib.rq_write_field(self._h, L.FIELDS[name], arr.ctypes.data, arr.nbytes))

    def _atoms_c(self):
        rows = [(self.atom_pos[n][0], self.atom_pos[n][1], self.atom_pos[n][2], self.Z_eff[n], self.atom_rc[n])
                for n in range(self.nAtoms)]
        return atoms_array(rows)

    def _chk(self, code):
        return L.check(self._lib, self._h, code)

    # ---- reference methods ----
    def initialize(self, U0=None):
        """initTrial, normalise, bootstrap Poisson (cloud/solver.py:402-419).

        ``U0`` (optional, S^3 floats): start amplitudes written over the trial
        U after the trial pass has set K and the labels (e.g. the seeded
        Gaussians of ``builders.gaussian_init``); they are normalised per
        domain before the bootstrap Poisson steps, as the trial U is."""
        if U0 is None:
            self._chk(self._lib.rq_initialize(self._h, self._bootstrap))
        else:
            self._chk(self._lib.rq_init_trial(self._h))
            self.U.set(U0)
            self._chk(self._lib.rq_normalize(self._h))
            self._chk(self._lib.rq_poisson(self._h, self._bootstrap))
        self._chk(self._lib.rq_sync(self._h))
        self._initialized = True

    def _normalize(self):
        self._chk(self._lib.rq_normalize(self._h))

    def _poisson_step(self):
        self._chk(self._lib.rq_poisson(self._h, 1))

    def _update(self):
        self._chk(self._lib.rq_update(self._h))

    def _boundary(self):
        self._chk(self._lib.rq_boundary(self._h))

    def _compute_energy(self):
        """T, V_eK, V_ee on the GPU (FP64 sums) + E_KK on the host (cloud/solver.py:446-476)."""
        out = (C.c_double * 6)()
        per = (C.c_double * (3 * self.nAtoms))()
        self._chk(self._lib.rq_energy(self._h, out, per))
        self.E_T, self.E_eK, self.E_ee = float(out[0]), float(out[1]), float(out[2])
        self.dipole = np.array([out[3], out[4], out[5]])
        self.E_domains = np.frombuffer(per, dtype=np.float64).reshape(self.nAtoms, 3).copy()
        self.E_KK = nuclear_energy(self.atom_pos, self.Z_eff, self.h)
        self.E = self.E_T + self.E_eK + self.E_ee + self.E_KK

    def _compute_forces(self):
        """Electronic HF forces (GPU, FP64 sums) + nucleus-nucleus (host) (cloud/solver.py:478-520)."""
        BATCH = 32
        nForce = self._nProtein
        # the reference's last batch ends at min(b0 + 32, nAtoms), not nForce
        count = 0
        for b0 in range(0, nForce, BATCH):
            count = min(b0 + BATCH, self.nAtoms)
        self.nuc_force = np.zeros((self.nAtoms, 3), dtype=np.float32)
        if count > 0:
            out = (C.c_double * (3 * count))()
            self._chk(self._lib.rq_forces_hf(self._h, 0, count, out))
            self.nuc_force[:count] = np.frombuffer(out, dtype=np.float64).reshape(count, 3)
        self.nuc_force[:nForce] += nuclear_forces(self.atom_pos, self.Z_eff, self.h, nForce)[:nForce].astype(np.float32)

    def _move_nuclei(self):
        """Damped nuclear step + K refresh (cloud/solver.py:522-548)."""
        mass_map = {1: 1 * 1836, 2: 16 * 1836, 3: 14 * 1836, 4: 12 * 1836}
        for n in range(self._nProtein):
            z = int(self.Z_eff[n])
            if z <= 0:
                continue
            m = mass_map.get(z, 1836)
            for d in range(3):
                self.nuc_vel[n][d] += self.nuc_force[n][d] / m * self.dt_nuc * self.force_scale
                self.nuc_vel[n][d] *= self.damping
                self.nuc_vel[n][d] = np.clip(self.nuc_vel[n][d], -self.max_vel, self.max_vel)
                self.atom_pos[n][d] += self.nuc_vel[n][d] * self.dt_nuc / self.h
                self.atom_pos[n][d] = np.clip(self.atom_pos[n][d], 5, self.NN - 5)
        self._chk(self._lib.rq_set_atoms(self._h, self._atoms_c(), self.nAtoms))
        self._chk(self._lib.rq_recompute_k(self._h))

    def run(self, total_steps, norm_interval=20, poisson_interval=2, force_interval=None, report_interval=500):
        """Run ``total_steps`` ITP steps; returns the list of snapshot dicts (cloud/solver.py:550-634).

        Steps between host events run as one rq_step call (no host round trip).
        Step numbering restarts at 1 on every call, as in the reference.
        """
        if force_interval is None:
            force_interval = self._config_force_interval
        if not self._initialized:
            self.initialize()
        sched = L.RqSchedule(int(poisson_interval), int(norm_interval), int(norm_interval * 5))
        results = []
        t0 = time.time()
        step = 0
        while step < total_steps:
            nxt = total_steps
            if report_interval:
                nxt = min(nxt, (step // report_interval + 1) * report_interval)
            if self.dynamics_enabled and force_interval:
                nxt = min(nxt, (step // force_interval + 1) * force_interval)
            try:
                self._chk(self._lib.rq_step(self._h, step, nxt - step, C.byref(sched)))
            except L.NumericalInstability as e:
                raise L.NumericalInstability(e.code, f"Numerical instability at step {nxt} ({e})") from None
            step = nxt
            if self.dynamics_enabled and force_interval and step % force_interval == 0:
                self._compute_forces()
                self._move_nuclei()
            if report_interval and step % report_interval == 0:
                self._compute_energy()
                elapsed = time.time() - t0
                ms_per_step = elapsed / step * 1000
                snapshot = {
                    'step': int(step),
                    'E': float(self.E),
                    'E_T': float(self.E_T),
                    'E_eK': float(self.E_eK),
                    'E_ee': float(self.E_ee),
                    'E_KK': float(self.E_KK),
                    'E_bind': float(self.E - self.E_atoms_sum),
                    'ms_per_step': float(ms_per_step),
                    'elapsed_sec': float(elapsed),
                    'nucPos': [[float(x) for x in row] for row in self.atom_pos],
                }
                if getattr(self, '_fold_atoms', None):
                    snapshot['fold_angle'] = self._fold_angle()
                results.append(snapshot)
                self._log_report(step, total_steps, ms_per_step)
        self._chk(self._lib.rq_sync(self._h))
        self._step += total_steps
        return results

    def set_fold_atoms(self, indices):
        self._fold_atoms = indices

    def _fold_angle(self):
        """Angle (degrees) at the middle fold atom between the other two,
        from atan2(|u x v|, u . v) of the bond vectors (0 if one is null)."""
        pa, pb, pc = (np.asarray(self.atom_pos[i], np.float64) * self.h for i in self._fold_atoms[:3])
        u, v = pa - pb, pc - pb
        if not (np.any(u) and np.any(v)):
            return 0.0
        return float(np.degrees(np.arctan2(np.linalg.norm(np.cross(u, v)), float(u @ v))))

    def _log_report(self, step, total, ms_per_step):
        """One progress line per report: energies in Ha, the largest force
        on a moving nucleus in Ha/bohr, the wall time per step so far."""
        moving = self.nuc_force[:self._nProtein]
        fmax = float(np.sqrt((moving.astype(np.float64) ** 2).sum(1)).max()) if len(moving) else 0.0
        print(f"[realqm] {step:>7d}/{total} steps  E {self.E:+.6f} Ha  "
              f"E-E_atoms {self.E - self.E_atoms_sum:+.6f} Ha  |F|max {fmax:.4f}  {ms_per_step:.3f} ms/step")

    # ---- extras ----
    def step(self, first_step, n_steps, poisson_interval=2, norm_interval=20, boundary_interval=None):
        """Run steps first_step+1 .. first_step+n_steps of the loop body on the device."""
        bi = boundary_interval if boundary_interval is not None else 5 * norm_interval
        sched = L.RqSchedule(int(poisson_interval), int(norm_interval), int(bi))
        self._chk(self._lib.rq_step(self._h, int(first_step), int(n_steps), C.byref(sched)))

    def pother_frame(self, frame):
        """End of a molecule.js frame with pother "frame": rebuild the
        frame potential (P minus the self-interaction terms on SIC frames,
        molecule.js:3009-3076); the next frame's ITP steps and the energy
        read it."""
        self._chk(self._lib.rq_pother_frame(self._h, int(frame)))

    def display_slice(self, slice_k=None):
        """molecule.js's visualisation slice at k = slice_k (default NN/2,
        molecule.js:1827): dict of the (S, S) density, label + 1000 Z and
        boundary planes and the K and cutoff-weight lines along i."""
        S = self.S
        k = self.NN // 2 if slice_k is None else int(slice_k)
        out = np.empty(3 * S * S + 2 * S, np.float32)
        self._chk(self._lib.rq_display_slice(self._h, k, out.ctypes.data_as(C.POINTER(C.c_float)), out.size))
        SS = S * S
        return {"density": out[:SS].reshape(S, S), "labelZ": out[SS:2 * SS].reshape(S, S),
                "boundary": out[2 * SS:3 * SS].reshape(S, S), "K_line": out[3 * SS:3 * SS + S],
                "w_line": out[3 * SS + S:], "raw": out}

    def domain_norms(self):
        out = (C.c_double * self.nAtoms)()
        self._chk(self._lib.rq_domain_norms(self._h, out))
        return np.frombuffer(out, dtype=np.float64).copy()

    def forces_gradp(self, radius=0):
        out = (C.c_double * (3 * self.nAtoms))()
        self._chk(self._lib.rq_forces_gradp(self._h, int(radius), out))
        return np.frombuffer(out, dtype=np.float64).reshape(self.nAtoms, 3).copy()

    def sync(self):
        self._chk(self._lib.rq_sync(self._h))

    def close(self):
        if self._h is not None:
            self._lib.rq_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def nuclear_energy(pos, Z, h, soft=1.0):
    """E_KK = sum_{a<b} Z_a Z_b / sqrt(R^2 + soft h^2), FP64 (cloud/solver.py:462-474)."""
    P = np.asarray(pos, np.float64) * h
    Zs = np.asarray(Z, np.float64)
    act = np.nonzero(Zs > 0)[0]
    if len(act) < 2:
        return 0.0
    P = P[act]
    Zs = Zs[act]
    d = P[:, None, :] - P[None, :, :]
    r = np.sqrt((d * d).sum(-1) + soft * h * h)
    iu = np.triu_indices(len(act), 1)
    return float((Zs[:, None] * Zs[None, :] / r)[iu].sum())


def nuclear_forces(pos, Z, h, n_force=None):
    """Force on each atom from the other nuclei, h^2-softened, FP64 (cloud/solver.py:504-520)."""
    P = np.asarray(pos, np.float64) * h
    Zs = np.asarray(Z, np.float64)
    n = len(Zs)
    n_force = n if n_force is None else n_force
    F = np.zeros((n, 3))
    act = Zs > 0
    d = P[:, None, :] - P[None, :, :]
    r2 = (d * d).sum(-1) + h * h
    w = (Zs[:, None] * Zs[None, :]) / (r2 * np.sqrt(r2))
    mask = act[:, None] & act[None, :]
    np.fill_diagonal(mask, False)
    w = np.where(mask, w, 0.0)
    F = (w[:, :, None] * d).sum(1)
    F[n_force:] = 0.0
    return F
