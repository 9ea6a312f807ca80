"""Host of the fused per-electron solver of mol_fast.js over librq.so.

``MolFast(config)`` takes the page's configuration globals as a dict (the
``window.USER_*`` names, mol_fast.js:13-93 and 815-826) and keeps the page's
host loop:

  do_steps(n)       doSteps(n)            mol_fast.js:1237-1324
  compute_ekk()     computeEKK            mol_fast.js:1326-1337
  compute_forces()  computeForces         mol_fast.js:1339-1373
  step_nuclei()     stepNuclei            mol_fast.js:1375-1410
  frame()           one draw() iteration  mol_fast.js:1439-1447

The device work (set-up, fused update, reductions, normalisation, 2-grid
corrections, forces, K rebuild, slice extract) runs in the HIP library
(include/rq_molfast.h); there is no CPU path.  Energies and dipoles are the
page's: E = E_T + E_eK + E_ee + E_KK from the reduction of the last step of
the frame.  Splits are given by name ('sphere', 'hemi', 'third', 'tetra',
'hemi_third'); the page also reads numeric split codes at set-up only, which
this host rejects.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib as L

PROTON_MASS = 1836


def _nuclei_table(raw):
    """The page's normalisation of USER_NUCLEI (mol_fast.js:29-46)."""
    out = []
    for n in raw:
        split = n.get("split") or "sphere"
        if not isinstance(split, str) or split not in L.RQ_MF_SPLIT:
            raise ValueError(f"mol_fast: unsupported split {split!r}")
        Z = float(n["Z"])
        out.append(dict(i=float(n["i"]), j=float(n["j"]), k=float(n["k"]), Z=Z, rc=float(n.get("rc") or 0),
                        r_out=float(n.get("r_out") or 0), tilt_y=float(n.get("tilt_y") or 0),
                        tilt_x=float(n.get("tilt_x") or 0), hs=float(n.get("hs") or 0),
                        fixed=n.get("fixed") is True,
                        z_init=float(n["z_init"]) if n.get("z_init") is not None else (Z if Z > 0 else 1.0),
                        split=split, split_idx=int(n.get("split_idx") or 0),
                        split_axis=tuple(n.get("split_axis") or (1, 0, 0)),
                        split_fix=bool(n["split_fix"]) if n.get("split_fix") is not None else True))
    return out


class MolFast:
    """mol_fast.js run state on one MI355X."""

    def __init__(self, config: dict):
        cfg = dict(config)
        self.NN = int(cfg.get("USER_NN") or 200)
        if self.NN % 2:
            raise ValueError("mol_fast: NN must be even for V-cycle")
        self.screen = float(cfg.get("USER_SCREEN") or 15)
        N2 = round(self.NN / 2)
        raw = cfg.get("USER_NUCLEI") or [dict(i=N2 - 20, j=N2, k=N2, Z=1, rc=0), dict(i=N2 + 20, j=N2, k=N2, Z=1, rc=0)]
        if len(raw) > L.RQ_MF_MAX_ELEC:
            raise ValueError(f"mol_fast: MAX_ATOMS = {L.RQ_MF_MAX_ELEC}, got {len(raw)}")
        self.nuclei = _nuclei_table(raw)
        self.NE = len(self.nuclei)
        self.targets = [float(t) for t in (cfg.get("USER_NORM_TARGETS") or [1] * self.NE)]
        self.S = self.NN + 1
        self.N2 = N2
        self.hv = self.screen / self.NN
        self.h2v = self.hv * self.hv
        self.steps_per_frame = int(cfg.get("USER_STEPS_PER_FRAME") or 500)
        self.max_steps = int(cfg.get("USER_STEPS") or 100000)
        self.r_out_release_step = cfg.get("USER_R_OUT_RELEASE_STEP")
        self.r_out_enforce = True
        self.dt_nuc = float(cfg.get("USER_DT_NUC") or 2.0)
        self.nuc_damping = float(cfg.get("USER_NUC_DAMPING") or 0.9)
        self.force_scale = float(cfg.get("USER_FORCE_SCALE") or 1.0)
        self.mass = list(cfg.get("USER_MASS") or [max(1, n["Z"]) * PROTON_MASS for n in self.nuclei])
        self.dynamics = bool(cfg.get("USER_DYNAMICS") or False)
        self.vel = [[0.0, 0.0, 0.0] for _ in range(self.NE)]
        self.force = [[0.0, 0.0, 0.0] for _ in range(self.NE)]
        self.t_step = 0
        self.E = self.E_T = self.E_eK = self.E_ee = self.E_KK = 0.0
        self.dipole = [0.0, 0.0, 0.0]
        self.sums = None

        p = L.RqMfParams()
        self._lib = L.load()
        self._lib.rq_mf_default_params(C.byref(p))
        p.NN = self.NN
        p.screen = self.screen
        p.shell_mode = 1 if cfg.get("USER_SHELL_MODE") is True else 0
        p.full_self = 1 if cfg.get("USER_FULL_SELF") else 0
        p.vcycle = 0 if cfg.get("USER_VCYCLE") is False else 1
        p.vcycle_interval = int(cfg.get("USER_VCYCLE_INTERVAL") or 50)
        p.coarse_iters = int(cfg.get("USER_COARSE_ITERS") or 10)
        p.kinetic_coeff = float(cfg["USER_KINETIC_COEFF"]) if cfg.get("USER_KINETIC_COEFF") is not None else 0.5
        p.kc_multi = float(cfg["USER_KC_MULTI"]) if cfg.get("USER_KC_MULTI") is not None else 1.0
        p.beta_pauli = float(cfg["USER_BETA_PAULI"]) if cfg.get("USER_BETA_PAULI") is not None else 0.0
        p.alpha_tf = float(cfg["USER_ALPHA_TF"]) if cfg.get("USER_ALPHA_TF") is not None else 0.0
        p.field_x = float(cfg["USER_FIELD_X"]) if cfg.get("USER_FIELD_X") is not None else 0.0
        self.params = p
        arr = (L.RqMfNucleus * self.NE)()
        for a, n in enumerate(self.nuclei):
            q = arr[a]
            q.i, q.j, q.k, q.Z = n["i"], n["j"], n["k"], n["Z"]
            q.rc, q.r_out, q.tilt_x, q.tilt_y, q.hs = n["rc"], n["r_out"], n["tilt_x"], n["tilt_y"], n["hs"]
            q.z_init, q.target = n["z_init"], self.targets[a]
            for d in range(3):
                q.axis[d] = float(n["split_axis"][d])
            q.split = L.RQ_MF_SPLIT[n["split"]]
            q.split_idx, q.split_fix = n["split_idx"], 1 if n["split_fix"] else 0
        h = C.c_void_p()
        self._check(self._lib.rq_mf_create(C.byref(p), arr, self.NE, C.byref(h)), None)
        self._h = h
        self._check(self._lib.rq_mf_init(self._h))
        self.compute_ekk()

    # ---- plumbing ----
    def _check(self, code, h=0):
        if code != L.RQ_OK:
            msg = self._lib.rq_mf_last_error(self._h if h == 0 else h)
            raise L.RqError(code, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.rq_mf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- fields ----
    def field(self, name: str, m: int = 0) -> np.ndarray:
        out = np.empty(self.S ** 3, np.float32)
        self._check(self._lib.rq_mf_read_field(self._h, L.RQ_MF_FIELDS[name], m,
                                               out.ctypes.data_as(C.POINTER(C.c_float)), out.size))
        return out.reshape(self.S, self.S, self.S)

    def set_field(self, name: str, m: int, data) -> None:
        a = np.ascontiguousarray(data, np.float32).ravel()
        self._check(self._lib.rq_mf_write_field(self._h, L.RQ_MF_FIELDS[name], m,
                                                a.ctypes.data_as(C.POINTER(C.c_float)), a.size))

    def extract(self) -> np.ndarray:
        out = np.empty(3 * self.NE * self.S * self.S, np.float32)
        self._check(self._lib.rq_mf_extract(self._h, out.ctypes.data_as(C.POINTER(C.c_float)), out.size))
        return out

    # ---- the page's host loop ----
    def do_steps(self, n: int):
        """doSteps(n): n fused steps on the device, then energies and dipole
        from the last reduction (mol_fast.js:1237-1324)."""
        if self.r_out_release_step is not None and self.r_out_enforce and self.t_step >= self.r_out_release_step:
            self.r_out_enforce = False
            self._check(self._lib.rq_mf_set_r_out(self._h, 0))
        sums = (C.c_double * (self.NE + 6))()
        self._check(self._lib.rq_mf_do_steps(self._h, int(n), sums))
        s = np.array(sums[:], np.float64)
        self.sums = s
        NE = self.NE
        self.E_T, self.E_eK, self.E_ee = float(s[NE]), float(s[NE + 1]), float(s[NE + 2])
        self.E = self.E_T + self.E_eK + self.E_ee + self.E_KK
        dn = [0.0, 0.0, 0.0]
        for n_ in self.nuclei:
            dn[0] += n_["Z"] * (n_["i"] - self.N2) * self.hv
            dn[1] += n_["Z"] * (n_["j"] - self.N2) * self.hv
            dn[2] += n_["Z"] * (n_["k"] - self.N2) * self.hv
        self.dipole = [dn[0] + float(s[NE + 3]), dn[1] + float(s[NE + 4]), dn[2] + float(s[NE + 5])]
        self.t_step += n
        if not math.isfinite(self.E):
            raise L.RqError(L.RQ_EUNSTABLE, f"numerical instability at step {self.t_step}")
        return self.E

    def compute_ekk(self):
        E = 0.0
        for a in range(self.NE):
            for b in range(a + 1, self.NE):
                na, nb = self.nuclei[a], self.nuclei[b]
                dx, dy, dz = (na["i"] - nb["i"]) * self.hv, (na["j"] - nb["j"]) * self.hv, (na["k"] - nb["k"]) * self.hv
                r = math.sqrt(dx * dx + dy * dy + dz * dz)
                if r > 0:
                    E += na["Z"] * nb["Z"] / r
        self.E_KK = E
        return E

    def compute_forces(self):
        fe = (C.c_float * (3 * self.NE))()
        self._check(self._lib.rq_mf_forces(self._h, fe))
        for a in range(self.NE):
            f = [float(fe[3 * a]), float(fe[3 * a + 1]), float(fe[3 * a + 2])]
            na = self.nuclei[a]
            for b in range(self.NE):
                if b == a:
                    continue
                nb = self.nuclei[b]
                dx, dy, dz = (na["i"] - nb["i"]) * self.hv, (na["j"] - nb["j"]) * self.hv, (na["k"] - nb["k"]) * self.hv
                r2 = dx * dx + dy * dy + dz * dz + self.h2v
                inv_r3 = 1 / (r2 * math.sqrt(r2))
                pref = na["Z"] * nb["Z"] * inv_r3
                f[0] += pref * dx
                f[1] += pref * dy
                f[2] += pref * dz
            self.force[a] = f
        return [list(f) for f in self.force]

    def step_nuclei(self):
        dt = self.dt_nuc
        moved = False
        for a, n in enumerate(self.nuclei):
            if n["Z"] <= 0 or n["fixed"]:
                continue
            m = self.mass[a]
            for d in range(3):
                self.vel[a][d] = self.nuc_damping * (self.vel[a][d] + dt * self.force_scale * self.force[a][d] / m)
            di, dj, dk = (self.vel[a][0] * dt) / self.hv, (self.vel[a][1] * dt) / self.hv, (self.vel[a][2] * dt) / self.hv
            if abs(di) > 0.0001 or abs(dj) > 0.0001 or abs(dk) > 0.0001:
                n["i"] += di
                n["j"] += dj
                n["k"] += dk
                moved = True
        if moved:
            pos = (C.c_double * (3 * self.NE))(*[x for n in self.nuclei for x in (n["i"], n["j"], n["k"])])
            self._check(self._lib.rq_mf_set_positions(self._h, pos))
            self.compute_ekk()
            self._check(self._lib.rq_mf_recompute_k(self._h))
        return moved

    def frame(self):
        """One draw() iteration: a frame of steps, forces, and a nuclear step
        when dynamics is on (mol_fast.js:1439-1447)."""
        if self.t_step >= self.max_steps:
            return self.E
        self.do_steps(self.steps_per_frame)
        self.compute_forces()
        if self.dynamics:
            self.step_nuclei()
        return self.E

    # ---- timing (bench) ----
    def set_timing(self, on: bool):
        self._check(self._lib.rq_mf_set_timing(self._h, 1 if on else 0))

    def update_time(self):
        ms = C.c_double()
        n = C.c_int32()
        self._check(self._lib.rq_mf_update_time(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value
