"""S2: CPU restatement of the molecule.js hot-path kernels (TEST INFRASTRUCTURE ONLY).

Same array conventions as ``oracle.s1``.  molecule.js stores nucleus
positions as u32 (``molecule.js:1882-1884``), so every position used here is
``floor(position)``.  Only the parts of the WGSL kernels that the build
implements are restated: the angular-sector ("split") constraints, bare
nuclei (Z = 0, Z_nuc > 0), the W-cutoff blend and the frame-stale
``PotherBuf`` are not.

Parity pinning (tests/test_oracle_s2.py):
  init_s2      scalar brute force per node (all nodes incl. faces)
  rc_mask      scalar brute force
  levelset     explicit-loop restatement + invariants (interior W >= 1,
               flips only to a neighbouring label at voxels with a foreign neighbour)
  energy_s2    explicit-loop restatement; dipole of a symmetric density
  gradp_forces linear P -> exact gradient (closed form); ball cell count
  pother_direct linearity: the per-domain potentials of a two-domain density
               sum to the total potential (same V-cycle, rhs split by label);
               the gathered P equals the other domain's potential
  self_potential / pother_frame  linearity (scaling U by sqrt 2 doubles the
               self-potential), additivity over domains (the self-potentials
               of all domains sum to the solve of the total density), Pother
               = P off SIC frames and changed only on the domain's own nodes
"""
from __future__ import annotations

import numpy as np

from .s1 import Grid, _interior, _nb

INV_SQRT_PI = 0.5641895835


def _fl(atoms):
    a = np.asarray(atoms, np.float64).reshape(-1, 5).copy()
    a[:, :3] = np.floor(a[:, :3])
    return a


def init_s2(atoms, g: Grid):
    """gpuInitAccum + gpuInitFinal (molecule.js:1377-1449) with initZeff = Z,
    initRcut = 1e6, Z_nuc = Z and no splits: K = sum Z/r~ with
    r~ = max(sqrt(r^2+0.04h^2), rc) (rc > 0) or max(r, 2h); label = first
    argmax of Z^2/sqrt(pi) exp(-Z r~); U = best trial -- on ALL nodes."""
    dt = g.dtype
    S = g.S
    atoms = _fl(atoms).astype(dt)
    x = np.arange(S).astype(dt) * g.h
    r_sing = dt(2.0 * g.h64)
    bU = np.zeros((S, S, S), dt)
    best = np.zeros((S, S, S), np.int32)
    K = np.zeros((S, S, S), dt)
    for a in range(atoms.shape[0]):
        ai, aj, ak, Za, rc = atoms[a]
        if not Za > 0:
            continue
        dx = (x - ai * g.h)[:, None, None]
        dy = (x - aj * g.h)[None, :, None]
        dz = (x - ak * g.h)[None, None, :]
        r2 = dx * dx + dy * dy + dz * dz
        if rc > 0:
            r = np.maximum(np.sqrt(r2 + dt(0.04) * g.h2), rc)
        else:
            r = np.maximum(np.sqrt(r2), r_sing)
        K = K + Za / r
        ut = Za * Za * dt(INV_SQRT_PI) * np.exp(-Za * r)
        upd = ut > bU
        bU = np.where(upd, ut, bU)
        best = np.where(upd, np.int32(a), best)
    return bU.copy(), best, K, bU


def rc_mask(label, atoms, g: Grid):
    """True where a node lies strictly inside rc of its own domain's nucleus
    (molecule.js:242-246 isInsideRc)."""
    dt = g.dtype
    S = g.S
    atoms = _fl(atoms).astype(dt)
    idx = np.arange(S).astype(dt)
    I, J, Kk = np.meshgrid(idx, idx, idx, indexing="ij")
    a = atoms[np.clip(label, 0, len(atoms) - 1)]
    dx = (I - a[..., 0]) * g.h
    dy = (J - a[..., 1]) * g.h
    dz = (Kk - a[..., 2]) * g.h
    r = np.sqrt(dx * dx + dy * dy + dz * dz)
    return (a[..., 4] > 0) & (r < a[..., 4])


def levelset(U, label, W, g: Grid, dt_w=0.5, curv_reg=0.15, lock_group=None):
    """One level-set step (molecule.js:426-540): returns (label', W').

    Per interior voxel: count same-label neighbours (order +i,-i,+j,-j,+k,-k)
    and take the foreign neighbour of largest U^2 (strict '>').  All six same:
    W = max(W, 1).  Locked pair: W = max(W, 0.5).  Otherwise
    W += dt_w (rho - rho_o)/(rho + rho_o) (0 if the sum < 1e-20) + curv_reg (cnt-3)/3,
    and W < 0 flips to the foreign label with W = 0.1.
    """
    dt = g.dtype
    S = g.S
    I = _interior(S)
    L = label[I]
    uc = U[I]
    my_rho = uc * uc
    cnt = np.zeros(L.shape, dt)
    best_rho = np.zeros(L.shape, dt)
    best_l = L.copy()
    for axis in range(3):
        for sign in (1, -1):
            ln = _nb(label, axis, sign)
            un = _nb(U, axis, sign)
            same = ln == L
            cnt = cnt + same.astype(dt)
            r = un * un
            upd = (~same) & (r > best_rho)
            best_rho = np.where(upd, r, best_rho)
            best_l = np.where(upd, ln, best_l)
    w = W[I].copy()
    out_l = L.copy()
    interior = cnt >= dt(6.0)
    locked = np.zeros(L.shape, bool)
    if lock_group is not None:
        lg = np.asarray(lock_group)
        gm = lg[np.clip(L, 0, len(lg) - 1)]
        gb = lg[np.clip(best_l, 0, len(lg) - 1)]
        locked = (~interior) & (gm > 0) & (gm == gb)
    denom = my_rho + best_rho
    with np.errstate(divide="ignore", invalid="ignore"):
        vel = np.where(denom < dt(1e-20), dt(0.0), (my_rho - best_rho) / denom)
    wn = w + dt(dt_w) * vel
    wn = wn + dt(curv_reg) * ((cnt - dt(3.0)) / dt(3.0))
    flip = wn < 0
    wn = np.where(flip, dt(0.1), wn)
    nl = np.where(flip, best_l, L)
    evolve = ~interior & ~locked
    w_out = np.where(interior, np.maximum(w, dt(1.0)), np.where(locked, np.maximum(w, dt(0.5)), wn))
    out_l = np.where(evolve, nl, L)
    Lo = label.copy()
    Wo = W.copy()
    Lo[I] = out_l
    Wo[I] = w_out
    return Lo, Wo


def energy_s2(U, K, P, label, atoms, g: Grid, n_domains=None):
    """reduceEnergy (molecule.js:950-1075): voxels inside max(rc_L, 2h) of their
    own nucleus are skipped; forward differences count toward same-label
    neighbours that are not inside that sphere; T, V_eK, V_ee per domain and
    the dipole sums sum rho x h^3 (x = i h) over the kept voxels.
    Returns (totals[3], per_domain[n,3], dipole[3])."""
    dt = g.dtype
    S = g.S
    I = _interior(S)
    atoms = _fl(atoms).astype(dt)
    n = atoms.shape[0] if n_domains is None else n_domains
    L = np.clip(label[I], 0, n - 1)
    a = atoms[L]
    r_ex = np.maximum(a[..., 4], dt(2.0 * g.h64))
    idx = np.arange(1, S - 1).astype(dt)

    def excl(oi, oj, ok):
        ii = (idx + oi)[:, None, None]
        jj = (idx + oj)[None, :, None]
        kk = (idx + ok)[None, None, :]
        dx = (ii - a[..., 0]) * g.h
        dy = (jj - a[..., 1]) * g.h
        dz = (kk - a[..., 2]) * g.h
        return np.sqrt(dx * dx + dy * dy + dz * dz) < r_ex

    keep = ~excl(0, 0, 0)
    u = U[I]
    rho = u * u
    d = []
    for axis, off in ((0, (1, 0, 0)), (1, (0, 1, 0)), (2, (0, 0, 1))):
        same = (np.clip(_nb(label, axis, 1), 0, n - 1) == L) & ~excl(*off)
        d.append(np.where(same, _nb(U, axis, 1) - u, dt(0.0)))
    T = dt(0.5) * (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) * g.h
    VeK = -K[I] * rho * g.h3
    Vee = P[I] * rho * g.h3
    per = np.zeros((n, 3))
    lk = L[keep]
    for c, t in enumerate((T, VeK, Vee)):
        per[:, c] = np.bincount(lk, weights=t[keep].astype(np.float64), minlength=n)
    xs = idx * g.h
    dip = []
    for axis in range(3):
        shape = [1, 1, 1]
        shape[axis] = -1
        coord = np.broadcast_to(xs.reshape(shape), rho.shape)
        dip.append((rho * coord * g.h3)[keep].astype(np.float64).sum())
    return per.sum(0), per, np.array(dip)


def gradp_forces(P, atoms, g: Grid, R=None):
    """gradPtotal (molecule.js:1199-1270): F = 2 Z <grad P> over integer cells
    within radius R of the (integer) nucleus, central differences, interior
    cells only; FP64 sums here (the reference sums in FP32 in one thread)."""
    dt = g.dtype
    S = g.S
    if R is None:
        R = max(5, int(round(3.0 / g.h64)))
    atoms = _fl(atoms)
    inv2h = dt(0.5) * (dt(1.0) / g.h)
    out = np.zeros((atoms.shape[0], 3))
    rr = np.arange(-R, R + 1)
    DI, DJ, DK = np.meshgrid(rr, rr, rr, indexing="ij")
    ball = DI * DI + DJ * DJ + DK * DK <= R * R
    for n, (ai, aj, ak, Z, _rc) in enumerate(atoms):
        if not Z > 0:
            continue
        ii = int(ai) + DI[ball]
        jj = int(aj) + DJ[ball]
        kk = int(ak) + DK[ball]
        ok = (ii >= 1) & (ii < S - 1) & (jj >= 1) & (jj < S - 1) & (kk >= 1) & (kk < S - 1)
        ii, jj, kk = ii[ok], jj[ok], kk[ok]
        if ii.size == 0:
            continue
        gi = (P[ii + 1, jj, kk] - P[ii - 1, jj, kk]) * inv2h
        gj = (P[ii, jj + 1, kk] - P[ii, jj - 1, kk]) * inv2h
        gk = (P[ii, jj, kk + 1] - P[ii, jj, kk - 1]) * inv2h
        cnt = float(ii.size)
        out[n] = [2.0 * Z * gi.astype(np.float64).sum() / cnt, 2.0 * Z * gj.astype(np.float64).sum() / cnt,
                  2.0 * Z * gk.astype(np.float64).sum() / cnt]
    return out


class S2State:
    """molecule.js frame schedule on the total-Poisson path (molecule.js:2822-3104):
    per step rho + ``jacobi`` damped sweeps (0.3333, 9), the two-grid correction
    when s > 0 and s % poisson_interval == 0, the rc-rule ITP update, normalise at
    (s+1) % 20 == 0 or the frame's last step; per frame the energy and
    ``ls_steps`` level-set steps.  FP32-faithful when dtype is float32.
    With ``pother_frame`` the ITP update and the energy read the frame
    potential (PotherBuf), rebuilt after each frame's steps by
    ``pother_frame`` with SIC every ``sic_interval`` frames (frames counted
    from 1, as the page's draw loop counts them)."""

    def __init__(self, rows, NN, screen, dtype=np.float32, dv=0.12, jacobi=2, poisson_interval=50,
                 curv_reg=0.15, ls_steps=None, pother_frame=False, sic_interval=0):
        from . import s1
        self.rows = np.asarray(rows, np.float64)
        self.g = s1.Grid(NN, screen, dv, dtype)
        self.U, self.label, self.K, _ = init_s2(self.rows, self.g)
        self.P = np.zeros_like(self.U)
        self.W = np.ones_like(self.U)
        self.jacobi = jacobi
        self.poisson_interval = poisson_interval
        self.curv_reg = curv_reg
        self.ls_steps = ls_steps
        self.use_frame = pother_frame
        self.sic_interval = sic_interval
        self.frame_no = 0
        self.Pother = self.P.copy()

    def frame(self, n):
        from . import s1, mg
        g = self.g
        dt = g.dtype
        Z = self.rows[:, 3]
        for s in range(n):
            rhs = s1.rho_rhs(self.U, g.two_pi, g.h2, dt)
            for _ in range(self.jacobi):
                self.P = s1.smooth(self.P, rhs, 0.3333, 9.0, dt)
            if s > 0 and s % self.poisson_interval == 0:
                self.P = mg.twogrid_s2(self.P, rhs, dt)
            mask = rc_mask(self.label, self.rows, g)
            Pv = self.Pother if self.use_frame else self.P
            self.U = s1.itp_update(self.U, self.K, Pv, self.label, g, rc_mask=mask)
            if (s + 1) % 20 == 0 or s == n - 1:
                self.U, _ = s1.normalize(self.U, self.label, Z, g)
        if self.use_frame:
            self.frame_no += 1
            self.Pother = pother_frame(self.U, self.label, self.P, Z, g, self.frame_no, self.sic_interval,
                                       0.3333, 9.0)
        Pv = self.Pother if self.use_frame else self.P
        tot, per, dip = energy_s2(self.U, self.K, Pv, self.label, self.rows, g)
        for _ in range(self.ls_steps if self.ls_steps is not None else max(1, n)):
            self.label, self.W = levelset(self.U, self.label, self.W, g, 0.5, self.curv_reg)
        return tot, per, dip


def rho_other(U, label, m, alpha, beta, dtype):
    """molecule.js:597-621 computeRhoOther in the V-cycle scaling: on the
    interior (alpha U^2) beta where label != m, 0 where label == m."""
    from .s1 import rho_rhs
    I = _interior(U.shape[0])
    r = rho_rhs(U, alpha, beta, dtype)
    return np.where((label[I] & 0x7FFF) != m, r, dtype(0)).astype(dtype)


def pother_direct(U, label, Pdom, Z, g: Grid, levels, nu1, nu2, a, c):
    """One direct-Pother Poisson step (molecule.js:2834-2858, with one
    multigrid V-cycle per domain in place of JACOBI_DIRECT sweeps): for every
    domain m with Z[m] > 0, P_m <- vcycle(P_m, rho_other_m); then
    P(x) = P_label(x)(x) on the interior (copyPotherForLabel, :664-688).
    Returns (P, Pdom)."""
    from . import mg
    dt = g.dtype
    S = g.S
    Pdom = [p.copy() for p in Pdom]
    for m in range(len(Pdom)):
        if not Z[m] > 0:
            continue
        rhs = rho_other(U, label, m, g.two_pi, g.h2, dt)
        Pdom[m] = mg.vcycle(Pdom[m], rhs, levels, nu1, nu2, a, c, dt)
    I = _interior(S)
    lab = (label[I] & 0x7FFF).astype(np.int64)
    P = np.zeros((S, S, S), dt)
    stack = np.stack([p[I] for p in Pdom])
    valid = lab < len(Pdom)
    P[I] = np.where(valid, np.take_along_axis(stack, np.minimum(lab, len(Pdom) - 1)[None], 0)[0], dt(0))
    return P, Pdom


def rho_self(U, label, m, alpha, beta, dtype):
    """molecule.js:572-593 computeRhoSelf in the V-cycle scaling: on the
    interior (alpha U^2) beta where label == m, else 0."""
    from .s1 import rho_rhs
    I = _interior(U.shape[0])
    r = rho_rhs(U, alpha, beta, dtype)
    return np.where((label[I] & 0x7FFF) == m, r, dtype(0)).astype(dtype)


def self_potential(U, label, m, g: Grid, a, c):
    """Domain m's self-potential solve of the SIC frame step
    (molecule.js:3018-3069): from zero, 4 damped sweeps, the sign-inverted
    two-grid correction (10 coarse sweeps, half the prolongation), 4 sweeps."""
    from . import mg
    dt = g.dtype
    rhs = rho_self(U, label, m, g.two_pi, g.h2, dt)
    x = np.zeros((g.S,) * 3, dt)
    for _ in range(4):
        x = mg.smooth(x, rhs, a, c, dt)
    x = mg.twogrid_s2(x, rhs, dt, a, c, 10, 0.5)
    for _ in range(4):
        x = mg.smooth(x, rhs, a, c, dt)
    return x


def pother_frame(U, label, P, Z, g: Grid, frame, sic_interval, a, c):
    """End-of-frame PotherBuf (molecule.js:3009-3076): a copy of P, minus on
    SIC frames (frame > 0, frame % sic_interval == 0) each domain's
    self-potential on its own nodes, scaled by 1/Z for Z > 1
    (subtractPself, :694-721)."""
    dt = g.dtype
    out = P.copy()
    if sic_interval > 0 and frame > 0 and frame % sic_interval == 0:
        I = _interior(g.S)
        lab = label[I] & 0x7FFF
        for m, z in enumerate(Z):
            z = dt(z)
            if z == 0:
                continue
            x = self_potential(U, label, m, g, a, c)
            frac = dt(dt(1.0) / z) if z > 1 else dt(1.0)
            out[I] = np.where(lab == m, out[I] - x[I] * frac, out[I])
    return out


def display_slice(U, label, K, atoms, sk, NN):
    """molecule.js:1108-1192 extract at k = sk with W cutoff 0, as the flat
    3 S^2 + 2 S float32 vector: density, label + 1000 Z, boundary marks,
    K line, cutoff-weight line (all 1)."""
    S = NN + 1
    f = np.float32
    atoms = np.asarray(atoms, np.float64).reshape(-1, 5)
    out = np.zeros(3 * S * S + 2 * S, f)
    u = U[:, :, sk].astype(f)
    lab = (np.asarray(label)[:, :, sk] & 0x7FFF).astype(np.int64)
    dens = u * u
    Zl = np.where(lab < len(atoms), atoms[np.minimum(lab, len(atoms) - 1), 3], 0.0).astype(f)
    inner = np.zeros((S, S), bool)
    inner[1:NN, 1:NN] = True
    out[:S * S] = np.where(inner, dens, f(0)).ravel()
    out[S * S:2 * S * S] = np.where(inner, lab.astype(f) + Zl * f(1000.0), f(0)).ravel()
    bnd = np.zeros((S, S), f)
    ii, jj = np.meshgrid(np.arange(S), np.arange(S), indexing="ij")
    # +i neighbour for 1 < i < NN-1
    lab_ip = np.zeros_like(lab)
    lab_ip[:-1] = lab[1:]
    d_ip = np.zeros_like(dens)
    d_ip[:-1] = dens[1:]
    ci = inner & (ii > 1) & (ii < NN - 1) & (lab != lab_ip)
    bnd = np.where(ci, np.where((dens > f(1e-6)) | (d_ip > f(1e-6)), f(1), f(2)), bnd)
    lab_jp = np.zeros_like(lab)
    lab_jp[:, :-1] = lab[:, 1:]
    d_jp = np.zeros_like(dens)
    d_jp[:, :-1] = dens[:, 1:]
    cj = inner & (bnd < f(1.5)) & (jj > 1) & (jj < NN - 1) & (lab != lab_jp)
    bnd = np.where(cj, np.where((dens > f(1e-6)) | (d_jp > f(1e-6)), np.maximum(bnd, f(1)), f(2)), bnd)
    out[2 * S * S:3 * S * S] = np.where(inner, bnd, f(0)).ravel()
    out[3 * S * S:3 * S * S + S] = K[:, sk, sk].astype(f)
    out[3 * S * S + S:] = f(1)
    return out
