"""S2: CPU restatement of the molecule.js hot-path kernels (TEST INFRASTRUCTURE ONLY).

Same array conventions as ``oracle.s1``.  molecule.js stores nucleus
positions as u32 (``molecule.js:1882-1884``), so every position used here is
``floor(position)``.  The per-atom fields of ``struct Atom`` beyond (i, j, k,
Z, rc) -- Z_nuc, initZeff, initRcut and the angular split (molecule.js:183-191,
1875-1903) -- come in as ``ext``, a list of dicts with the keys of
``rq_atom_s2`` (``resolve_ext`` applies the page's defaults).  The W-cutoff
blend (off by default) is not restated.

Parity pinning (tests/test_oracle_s2.py):
  init_s2      scalar brute force per node (all nodes incl. faces), with
               Z_nuc / initZeff / initRcut / splits too
  in_split     the hemisphere by the sign of the axis projection, a 120-degree
               rotation about the axis moves a node to the next third
  recompute_k_s2  closed form on a single atom; equals init K for rc = 0
               atoms away from the 2h clamp
  nuclear_energy_unique  merging co-located nuclei (closed form)
  pdirect_init / direct_sweeps_step  the start potential is the others'
               0.5/r sum (scalar loop); sweeps are s1.smooth on rho_other
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


SPLIT_CODE = {"sphere": 0, "hemi": 2, "third": 3}


def resolve_ext(rows, ext=None):
    """Per-atom S2 fields with molecule.js's defaults (fillAtomBuf,
    molecule.js:1875-1903): Zn = Z_nuc || Z, zeff = initZeff || Z, rcut =
    initRcut || 1e6, split sphere, axis (0, 0, 1)."""
    rows = np.asarray(rows, np.float64).reshape(-1, 5)
    out = []
    for n in range(rows.shape[0]):
        e = (ext[n] if ext is not None else None) or {}
        Z = rows[n, 3]
        st = e.get("split_type", e.get("split", 0))
        out.append({"Zn": e.get("Z_nuc", 0.0) if e.get("Z_nuc", 0.0) > 0 else Z,
                    "zeff": e.get("init_zeff", 0.0) if e.get("init_zeff", 0.0) > 0 else Z,
                    "rcut": e.get("init_rcut", 0.0) if e.get("init_rcut", 0.0) > 0 else 1e6,
                    "st": SPLIT_CODE[st] if isinstance(st, str) else int(st),
                    "sidx": int(e.get("split_idx", 0)),
                    "axis": tuple(e.get("split_axis", (0.0, 0.0, 1.0))),
                    "rot": float(e.get("split_rot", 0.0))})
    return out


def in_split(dx, dy, dz, r, e, dt):
    """inSplitAtom (molecule.js:194-225) on arrays of node offsets d (bohr)
    and |d|, FP32 operation order."""
    st, sidx = e["st"], e["sidx"]
    full = np.broadcast_shapes(np.shape(dx), np.shape(dy), np.shape(dz), np.shape(r))
    if st <= 1:
        return np.ones(full, bool)
    ax, ay, az = (dt(v) for v in e["axis"])
    axL = max(np.sqrt(ax * ax + ay * ay + az * az), dt(1e-12))
    a0, a1, a2 = ax / axL, ay / axL, az / axL
    dotp = dx * a0 + dy * a1 + dz * a2
    small = np.broadcast_to(r < dt(1e-6), full)
    if st == 2:
        res = (dotp < 0) if sidx == 0 else (dotp >= 0)
    elif st == 3:
        px, py, pz = dx - dotp * a0, dy - dotp * a1, dz - dotp * a2
        if abs(a2) < dt(0.9):
            e1x, e1y, e1z = a1, -a0, dt(0.0)
        else:
            e1x, e1y, e1z = dt(0.0), a2, -a1
        e1L = max(np.sqrt(e1x * e1x + e1y * e1y + e1z * e1z), dt(1e-12))
        e1x, e1y, e1z = e1x / e1L, e1y / e1L, e1z / e1L
        e2x, e2y, e2z = a1 * e1z - a2 * e1y, a2 * e1x - a0 * e1z, a0 * e1y - a1 * e1x
        pe1 = px * e1x + py * e1y + pz * e1z
        pe2 = px * e2x + py * e2y + pz * e2z
        th = np.arctan2(pe2, pe1).astype(dt) - dt(e["rot"])
        th = np.where(th < dt(-3.14159265), th + dt(6.28318530), th)
        th = np.where(th > dt(3.14159265), th - dt(6.28318530), th)
        sec = np.maximum(np.floor((th + dt(3.14159265)) / dt(2.09439510)), 0).astype(np.int64) % 3
        res = sec == sidx
    else:
        res = np.ones(full, bool)
    return small | np.broadcast_to(res, full)


def init_s2(atoms, g: Grid, ext=None):
    """gpuInitAccum + gpuInitFinal (molecule.js:1377-1449) on ALL nodes: K =
    sum Zn/r~ over atoms with Z > 0 or Zn > 0 (Zn = Z_nuc || Z), r~ =
    max(sqrt(r^2+0.04h^2), rc) (rc > 0) or max(r, 2h); the trial
    zeff^2/sqrt(pi) exp(-zeff r~) of atoms with Z > 0, zero where the real
    distance exceeds initRcut or the node lies outside the atom's split
    sector; label = first argmax of the trial; U = best trial."""
    dt = g.dtype
    S = g.S
    ex = resolve_ext(atoms, ext)
    atoms = _fl(atoms).astype(dt)
    x = np.arange(S).astype(dt) * g.h
    r_sing = dt(2.0 * g.h64)
    bU = np.zeros((S, S, S), dt)
    best = np.zeros((S, S, S), np.int32)
    K = np.zeros((S, S, S), dt)
    for a in range(atoms.shape[0]):
        ai, aj, ak, Za, rc = atoms[a]
        e = ex[a]
        Zn = dt(e["Zn"])
        if Za <= 0 and Zn <= 0:
            continue
        dx = np.broadcast_to((x - ai * g.h)[:, None, None], (S, S, S))
        dy = np.broadcast_to((x - aj * g.h)[None, :, None], (S, S, S))
        dz = np.broadcast_to((x - ak * g.h)[None, None, :], (S, S, S))
        r2 = dx * dx + dy * dy + dz * dz
        if rc > 0:
            r = np.maximum(np.sqrt(r2 + dt(0.04) * g.h2), rc)
        else:
            r = np.maximum(np.sqrt(r2), r_sing)
        K = K + Zn / r
        Ze = dt(e["zeff"])
        r_real = np.sqrt(r2)
        ok = ~(r_real > dt(e["rcut"])) & bool(Za > 0) & in_split(dx, dy, dz, r_real, e, dt)
        ut = np.where(ok, Ze * Ze * dt(INV_SQRT_PI) * np.exp(-Ze * r), dt(0.0)).astype(dt)
        upd = ut > bU
        bU = np.where(upd, ut, bU)
        best = np.where(upd, np.int32(a), best)
    return bU.copy(), best, K, bU


def recompute_k_s2(atoms, g: Grid, ext=None):
    """recomputeK (molecule.js:1337-1366) on all nodes: offsets (i - posI) h,
    K = sum Zn / max(r, rc) (rc > 0) or Zn / max(r, 2h)."""
    dt = g.dtype
    S = g.S
    ex = resolve_ext(atoms, ext)
    atoms = _fl(atoms).astype(dt)
    idx = np.arange(S).astype(dt)
    r_sing = dt(2.0 * g.h64)
    K = np.zeros((S, S, S), dt)
    for a in range(atoms.shape[0]):
        ai, aj, ak, Za, rc = atoms[a]
        Zn = dt(ex[a]["Zn"])
        if Za <= 0 and Zn <= 0:
            continue
        di = ((idx - ai) * g.h)[:, None, None]
        dj = ((idx - aj) * g.h)[None, :, None]
        dk = ((idx - ak) * g.h)[None, None, :]
        r = np.sqrt(di * di + dj * dj + dk * dk)
        r = np.maximum(r, rc) if rc > 0 else np.maximum(r, r_sing)
        K = K + Zn / r
    return K


def nuclear_energy_unique(pos, Z, Z_nuc, h):
    """E_KK of molecule.js (:55-62, 3166-3180): electrons at one position
    share one nucleus of charge sum(Z_nuc || Z); pairs of distinct nuclei,
    softened by 0.04 h^2.  FP64."""
    nuc = {}
    order = []
    for e in range(len(Z)):
        zn = Z_nuc[e] if Z_nuc is not None else 0.0
        if Z[e] == 0 and zn == 0:
            continue
        key = tuple(pos[e])
        if key not in nuc:
            nuc[key] = 0.0
            order.append(key)
        nuc[key] += zn or Z[e]
    e = 0.0
    for a in range(len(order)):
        for b in range(a + 1, len(order)):
            d2 = sum(((order[a][c] - order[b][c]) * h) ** 2 for c in range(3))
            e += nuc[order[a]] * nuc[order[b]] / np.sqrt(d2 + 0.04 * h * h)
    return e


def pdirect_init(rows, g: Grid, m):
    """initPdirect (molecule.js:624-662): sum over the other atoms with Z > 0
    of 0.5/sqrt(r^2 + h^2) on the interior, u32 positions, atom order."""
    dt = g.dtype
    S = g.S
    atoms = _fl(rows).astype(dt)
    x = np.arange(S).astype(dt) * g.h
    out = np.zeros((S, S, S), dt)
    I = _interior(S)
    val = np.zeros((S - 2,) * 3, dt)
    for n in range(atoms.shape[0]):
        if atoms[n, 3] <= 0 or n == m:
            continue
        dx = (x[1:-1] - atoms[n, 0] * g.h)[:, None, None]
        dy = (x[1:-1] - atoms[n, 1] * g.h)[None, :, None]
        dz = (x[1:-1] - atoms[n, 2] * g.h)[None, None, :]
        val = val + dt(0.5) / np.sqrt(dx * dx + dy * dy + dz * dz + g.h2)
    out[I] = val
    return out


def direct_sweeps_setup(rows, g: Grid):
    """The start of molecule.js's direct Pother: (Pdom, Pother) with Pother
    the sum of every charged domain's P_m (molecule.js:2678-2693)."""
    Z = np.asarray(rows, np.float64)[:, 3]
    Pdom = []
    Pother = np.zeros((g.S,) * 3, g.dtype)
    for m in range(len(Z)):
        if Z[m] == 0:
            Pdom.append(np.zeros((g.S,) * 3, g.dtype))
            continue
        Pm = pdirect_init(rows, g, m)
        Pdom.append(Pm)
        Pother = Pother + Pm
    return Pdom, Pother


def direct_sweeps_step(U, label, Pdom, Pother, Z, g: Grid, a=0.3333, c=9.0, sweeps=10):
    """One step of molecule.js's direct Pother (molecule.js:2833-2858): for
    each domain m with Z != 0, rho_other_m scaled (h^2 2pi) rho, `sweeps`
    damped sweeps of P_m, and Pother = P_m on the nodes labelled m."""
    from .s1 import smooth
    dt = g.dtype
    I = _interior(g.S)
    Pdom = [p.copy() for p in Pdom]
    Pother = Pother.copy()
    alpha = g.h2 * g.two_pi
    for m in range(len(Z)):
        if Z[m] == 0:
            continue
        rhs = rho_other(U, label, m, alpha, dt(1.0), dt)
        for _ in range(sweeps):
            Pdom[m] = smooth(Pdom[m], rhs, a, c, dt)
        sel = (label[I] & 0x7FFF) == m
        Pother[I] = np.where(sel, Pdom[m][I], Pother[I])
    return Pdom, Pother


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


def _sector_ok(L, atoms_fl, ex, g, dt):
    """cellSectorOK (molecule.js:438-443) of every interior node for the
    domain in L (same shape as the interior)."""
    S = g.S
    idx = np.arange(1, S - 1).astype(dt)
    ok = np.ones(L.shape, bool)
    for n, e in enumerate(ex):
        if e["st"] <= 1:
            continue
        sel = L == n
        if not sel.any():
            continue
        di = ((idx - dt(atoms_fl[n, 0])) * g.h)[:, None, None]
        dj = ((idx - dt(atoms_fl[n, 1])) * g.h)[None, :, None]
        dk = ((idx - dt(atoms_fl[n, 2])) * g.h)[None, None, :]
        di, dj, dk = np.broadcast_arrays(di, dj, dk)
        r = np.sqrt(di * di + dj * dj + dk * dk)
        ok = np.where(sel, in_split(di, dj, dk, r, e, dt), ok)
    return ok


def levelset(U, label, W, g: Grid, dt_w=0.5, curv_reg=0.15, lock_group=None, atoms=None, ext=None):
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
    if ext is not None and any(e["st"] > 1 for e in resolve_ext(atoms, ext)):
        # split siblings own only the nodes of their sector (molecule.js:521-536)
        ex = resolve_ext(atoms, ext)
        afl = _fl(atoms)
        my_ok = _sector_ok(L, afl, ex, g, dt)
        new_ok = _sector_ok(nl, afl, ex, g, dt)
        best_ok = _sector_ok(best_l, afl, ex, g, dt)
        cancel = ~new_ok
        rescue = cancel & ~my_ok & (best_l != L) & best_ok
        release = new_ok & ~my_ok & (nl == L) & (best_l != L) & best_ok
        nl = np.where(cancel, L, nl)
        wn = np.where(cancel, np.maximum(wn, dt(0.1)), wn)
        nl = np.where(rescue | release, best_l, nl)
        wn = np.where(rescue | release, dt(0.1), wn)
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


def gradp_forces(P, atoms, g: Grid, R=None, ext=None):
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
    ex = resolve_ext(atoms, ext)
    for n, (ai, aj, ak, Z, _rc) in enumerate(atoms):
        if Z <= 0:  # a bare nucleus feels the force with its nuclear charge (molecule.js:1214)
            Z = ex[n]["Zn"]
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
                 curv_reg=0.15, ls_steps=None, pother_frame=False, sic_interval=0, direct=False, ext=None):
        from . import s1
        self.rows = np.asarray(rows, np.float64)
        self.ext = ext
        self.g = s1.Grid(NN, screen, dv, dtype)
        self.U, self.label, self.K, _ = init_s2(self.rows, self.g, ext)
        self.P = np.zeros_like(self.U)
        # molecule.js's direct Pother (its default up to 5 electrons)
        self.direct = direct
        if direct:
            self.Pdom, self.Pother_direct = direct_sweeps_setup(self.rows, self.g)
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
            if self.direct:
                self.Pdom, self.Pother_direct = direct_sweeps_step(self.U, self.label, self.Pdom,
                                                                   self.Pother_direct, Z, g)
            else:
                rhs = s1.rho_rhs(self.U, g.two_pi, g.h2, dt)
                for _ in range(self.jacobi):
                    self.P = s1.smooth(self.P, rhs, 0.3333, 9.0, dt)
                if s > 0 and s % self.poisson_interval == 0:
                    self.P = mg.twogrid_s2(self.P, rhs, dt)
            mask = rc_mask(self.label, self.rows, g)
            Pv = self.Pother_direct if self.direct else (self.Pother if self.use_frame else self.P)
            self.U = s1.itp_update(self.U, self.K, Pv, self.label, g, rc_mask=mask)
            if (s + 1) % 20 == 0 or s == n - 1:
                self.U, _ = s1.normalize(self.U, self.label, Z, g)
        if self.use_frame and not self.direct:
            self.frame_no += 1
            self.Pother = pother_frame(self.U, self.label, self.P, Z, g, self.frame_no, self.sic_interval,
                                       0.3333, 9.0)
        Pv = self.Pother_direct if self.direct else (self.Pother if self.use_frame else self.P)
        tot, per, dip = energy_s2(self.U, self.K, Pv, self.label, self.rows, g)
        for _ in range(self.ls_steps if self.ls_steps is not None else max(1, n)):
            self.label, self.W = levelset(self.U, self.label, self.W, g, 0.5, self.curv_reg, atoms=self.rows,
                                          ext=self.ext)
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
