"""mol_fast.js: CPU restatement of the fused per-electron solver (TEST INFRASTRUCTURE ONLY).

Only tests/, the smoke check and the benchmark's CPU baseline use this module,
as the checker of the HIP path; the library never imports it.

mol_fast.js keeps one amplitude u_m, one partition weight w_m and one
Hartree potential P_m per electron (NELEC = number of nuclei), all on the
same S^3 grid, and advances them together in one fused update
(mol_fast.js:134-351).  Every field here is a float32 array of shape
(NELEC, S, S, S) (or (S, S, S) for K), node (i, j, k) at [.., i, j, k], so
the flat index i*S*S + j*S + k of the reference is the C-order index.

The WGSL passes are FP32; their expressions are restated with numpy float32
arithmetic in the order WGSL evaluates them (left-associative chains, no
fused multiply-add).  The host-side set-up (uploadInitialData,
mol_fast.js:941-1046) is JavaScript double precision and is restated in
float64, rounded to float32 where the page stores to a Float32Array.
Reductions: the page sums per-node FP32 terms in an FP32 tree; here the same
per-node FP32 terms are summed in float64.

Parity pinning (tests/test_oracle_molfast.py):
  consts          closed form of writeParams (mol_fast.js:838-876)
  in_split        sector membership brute force: hemi halves by sign, thirds
                  are 120-degree wedges (rotating a point by 120 degrees about
                  the axis moves it to the next sector), tetra by nearest vertex
  init            scalar per-node loop restatement at sampled nodes; the
                  pre-normalised U has norm = target; w partition disjoint
  update          explicit per-node scalar loop at sampled nodes; a uniform
                  field (u, w, P constant) gives the closed-form increments;
                  P-part linearity in the source
  reduce          closed forms: norm of a constant field, T of a linear ramp,
                  dipole of a symmetric density
  twogrid         restriction of a constant (full weighting sums to 1),
                  prolongation of a constant, coarse smoother fixed point
  forces          linear P gives the exact gradient times 2Z
  recompute_k     equals the init K formula in float32 on a one-atom grid
  extract         scalar loop restatement
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

f32 = np.float32

# split type codes of the update shader (mol_fast.js:118-119)
SPLIT_CODE = {"sphere": 0, "hemi": 2, "third": 3, "tetra": 4, "hemi_third": 5}
R_INIT = 3.0  # au, initial w extent (mol_fast.js:947)
FORCE_RADIUS = 4  # cells (mol_fast.js:494)


def _fixed6(x):
    """A JS number templated into WGSL with toFixed(6), then read as f32."""
    return f32(float(f"{float(x):.6f}"))


@dataclass
class Nucleus:
    """One entry of window.nuclei (mol_fast.js:29-46).  split_rot and r_in
    are not copied by that map, so the shader and the initialiser see 0."""
    i: float
    j: float
    k: float
    Z: float
    rc: float = 0.0
    r_out: float = 0.0
    tilt_y: float = 0.0
    tilt_x: float = 0.0
    hs: float = 0.0
    z_init: float | None = None
    split: object = "sphere"
    split_idx: int = 0
    split_axis: tuple = (1.0, 0.0, 0.0)
    split_fix: bool = True
    fixed: bool = False

    def __post_init__(self):
        if self.z_init is None:
            self.z_init = self.Z if self.Z > 0 else 1.0


@dataclass
class MolFast:
    """Static tables and scalars of one mol_fast.js run."""
    nuclei: list
    NN: int = 200
    screen: float = 15.0
    targets: list | None = None
    kinetic_coeff: float = 0.5
    kc_multi: float = 1.0
    shell_mode: bool = False
    beta_pauli: float = 0.0
    alpha_tf: float = 0.0
    full_self: bool = False
    field_x: float = 0.0
    r_out_enforce: bool = True
    vcycle: bool = True
    vcycle_interval: int = 50
    coarse_iters: int = 10
    norm_interval: int = 20
    extra: dict = field(default_factory=dict)

    def __post_init__(self):
        if self.NN % 2:
            raise ValueError("mol_fast: NN must be even for V-cycle")
        if self.targets is None:
            self.targets = [1.0] * len(self.nuclei)

    # ---- writeParams (mol_fast.js:838-876) ----
    @property
    def S(self):
        return self.NN + 1

    @property
    def NE(self):
        return len(self.nuclei)

    def consts(self):
        hv = self.screen / self.NN
        h2v = hv * hv
        h3v = hv * hv * hv
        dv = 0.12
        return dict(h=f32(hv), h2=f32(h2v), inv_h=f32(1.0 / hv), inv_h2=f32(1.0 / h2v), dt=f32(dv * h2v),
                    half_d=f32(self.kinetic_coeff * dv), h3=f32(h3v), two_pi=f32(2.0 * math.pi),
                    R_out=f32(1.0 if self.r_out_enforce else 0.0), field_x=f32(self.field_x),
                    alpha_tf=f32(self.alpha_tf), full_self=f32(1.0 if self.full_self else 0.0),
                    N2=int(round(self.NN / 2)), kc=_fixed6(self.kinetic_coeff), kc_multi=_fixed6(self.kc_multi),
                    beta=_fixed6(self.beta_pauli))

    def atoms_f32(self):
        return [(f32(n.i), f32(n.j), f32(n.k), f32(n.Z)) for n in self.nuclei]

    # ---- templated constant arrays of the update shader (mol_fast.js:96-131) ----
    def orbital_tables(self):
        tab = []
        for idx, n in enumerate(self.nuclei):
            ax = n.split_axis or (1.0, 0.0, 0.0)
            L = math.hypot(*ax) or 1.0
            axis = tuple(_fixed6(v / L) for v in ax)
            code = SPLIT_CODE.get(n.split, 0) if isinstance(n.split, str) else 0
            group = idx
            if n.split and n.split != "sphere":
                for j in range(idx):
                    m = self.nuclei[j]
                    if m.split == n.split and m.i == n.i and m.j == n.j and m.k == n.k:
                        group = j
                        break
            tab.append(dict(r_out=_fixed6(n.r_out or 0.0), split=code, sidx=int(n.split_idx or 0), axis=axis,
                            rot=f32(0.0), fix=bool(n.split_fix), group=group, rc=f32(n.rc),
                            target=f32(self.targets[idx]), tilt_x=f32(n.tilt_x or 0.0),
                            tilt_y=f32(n.tilt_y or 0.0)))
        return tab


# ---------------------------------------------------------------------------
# host-side set-up (JavaScript doubles)
# ---------------------------------------------------------------------------
def in_split(dx, dy, dz, r, split, split_idx, split_axis, split_rot=0.0):
    """inSplit (mol_fast.js:878-939), float64, vectorised over nodes."""
    dx, dy, dz, r = (np.asarray(a, np.float64) for a in (dx, dy, dz, r))
    true = np.ones(np.broadcast(dx, dy, dz, r).shape, bool)
    if split == "sphere" or split == 1:
        return true
    ax = split_axis or (1.0, 0.0, 0.0)
    L = math.hypot(ax[0], ax[1], ax[2]) or 1.0
    a0, a1, a2 = ax[0] / L, ax[1] / L, ax[2] / L
    dot = dx * a0 + dy * a1 + dz * a2

    def sector():
        px, py, pz = dx - dot * a0, dy - dot * a1, dz - dot * a2
        if abs(a2) < 0.9:
            e1 = [a1, -a0, 0.0]
        else:
            e1 = [0.0, a2, -a1]
        e1l = math.hypot(*e1) or 1.0
        e1x, e1y, e1z = e1[0] / e1l, e1[1] / e1l, e1[2] / e1l
        e2x, e2y, e2z = a1 * e1z - a2 * e1y, a2 * e1x - a0 * e1z, a0 * e1y - a1 * e1x
        pe1 = px * e1x + py * e1y + pz * e1z
        pe2 = px * e2x + py * e2y + pz * e2z
        theta = np.arctan2(pe2, pe1) - (split_rot or 0.0)
        theta = np.where(theta < -math.pi, theta + 2 * math.pi, theta)
        theta = np.where(theta > math.pi, theta - 2 * math.pi, theta)
        return np.floor((theta + math.pi) / (2 * math.pi / 3)).astype(np.int64) % 3

    if split in ("hemi", 2):
        out = (dot < 0) if split_idx == 0 else (dot >= 0)
    elif split in ("third", 3):
        out = sector() == split_idx
    elif split in ("tetra", 4):
        rr = np.where(r > 0, r, 1.0)
        ur = (dx / rr, dy / rr, dz / rr)
        best_v = np.zeros(true.shape, np.int64)
        best_d = np.full(true.shape, -np.inf)
        for iv, v in enumerate(((1, 1, 1), (1, -1, -1), (-1, 1, -1), (-1, -1, 1))):
            d = v[0] * ur[0] + v[1] * ur[1] + v[2] * ur[2]
            better = d > best_d
            best_d = np.where(better, d, best_d)
            best_v = np.where(better, iv, best_v)
        out = best_v == split_idx
    elif split in ("hemi_third", 5):
        if split_idx == 0:
            out = dot >= 0
        else:
            out = (dot < 0) & (sector() == split_idx - 1)
    else:
        return true
    return out | (r < 1e-6)


def init_fields(mf: MolFast):
    """uploadInitialData (mol_fast.js:941-1046): K, U, W, P on all nodes,
    float64 then float32, U pre-normalised to its target."""
    S, NN = mf.S, mf.NN
    hv = mf.screen / NN
    h2v = hv * hv
    h3v = hv * hv * hv
    soft = 4 * h2v
    ax = np.arange(S, dtype=np.float64)
    I, J, Kk = np.meshgrid(ax, ax, ax, indexing="ij")
    K = np.zeros((S, S, S))
    best_u = np.full((S, S, S), -1.0)
    best_m = np.full((S, S, S), -1, np.int64)
    r_list, u_list = [], []
    nuc = mf.nuclei
    mid_i0 = (nuc[0].i + nuc[-1].i) / 2
    for a, n in enumerate(nuc):
        dx, dy, dz = (I - n.i) * hv, (J - n.j) * hv, (Kk - n.k) * hv
        r2 = dx * dx + dy * dy + dz * dz
        r_soft = np.sqrt(r2 + soft)
        r = np.sqrt(r2)
        K = K + n.Z / r_soft
        u = np.exp(-n.z_init * r)
        r_list.append(r)
        u_list.append(u)
        tgt = mf.targets[a] or 1
        uw = math.sqrt(tgt * n.z_init * n.z_init * n.z_init) * u
        roa = n.r_out or 0
        elig_r = ~((roa > 0) & (r > roa + hv))
        hsa = n.hs or 0
        elig_hs = True if hsa == 0 else ((hsa > 0) == (I > mid_i0))
        elig_s = in_split(dx, dy, dz, r, n.split, n.split_idx, n.split_axis, 0.0)
        take = elig_r & elig_hs & elig_s & (uw > best_u)
        best_u = np.where(take, uw, best_u)
        best_m = np.where(take, a, best_m)

    def smoothcut(r, rc):
        edge = rc - 3 * hv
        t = np.maximum(0.0, np.minimum(1.0, (r - edge) / (rc - edge)))
        return t * t * (3 - 2 * t)

    NE = len(nuc)
    U = np.zeros((NE, S, S, S), np.float32)
    W = np.zeros((NE, S, S, S), np.float32)
    P = np.zeros((NE, S, S, S), np.float32)
    for m, n in enumerate(nuc):
        u = u_list[m]
        if n.rc > 0:
            u = np.where(r_list[m] < n.rc, u * smoothcut(r_list[m], n.rc), u)
        U[m] = u.astype(np.float32)
        P[m] = (0.5 * n.Z / (r_list[m] + h2v)).astype(np.float32)
    r_best = np.zeros((S, S, S))
    for a in range(NE):
        r_best = np.where(best_m == a, r_list[a], r_best)
    for a, n in enumerate(nuc):
        sel = (best_m == a) & (r_best < R_INIT)
        w = np.ones((S, S, S))
        if n.rc > 0:
            w = np.where(r_best < n.rc, smoothcut(r_best, n.rc), w)
        ro = n.r_out or 0
        if ro > 0:
            w = np.where(r_best > ro + hv, 0.0, w)
        W[a] = np.where(sel, w, 0.0).astype(np.float32)
    for m in range(NE):
        norm = float(np.sum(U[m].astype(np.float64) ** 2)) * h3v
        tgt = mf.targets[m]
        if norm > 0 and tgt > 0:
            U[m] = (U[m].astype(np.float64) * math.sqrt(tgt / norm)).astype(np.float32)
    return K.astype(np.float32), U, W, P


# ---------------------------------------------------------------------------
# FP32 passes
# ---------------------------------------------------------------------------
def _interior(S):
    return (slice(1, S - 1),) * 3


def _nb(a, di, dj, dk):
    S = a.shape[-1]
    return a[..., 1 + di:S - 1 + di, 1 + dj:S - 1 + dj, 1 + dk:S - 1 + dk]


def _coords(S):
    ax = np.arange(1, S - 1, dtype=np.float32)
    return np.meshgrid(ax, ax, ax, indexing="ij")


def _sector_mask_f32(m, tab, rel, r_m):
    """Angular split test of the update shader (mol_fast.js:194-263), FP32.
    normalize(v) is taken as v / sqrt(dot(v, v))."""
    s_type, sidx = tab["split"], tab["sidx"]
    axv = tab["axis"]
    rx, ry, rz = rel
    dotA = (rx * axv[0] + ry * axv[1]) + rz * axv[2]

    def sector():
        ppx, ppy, ppz = rx - dotA * axv[0], ry - dotA * axv[1], rz - dotA * axv[2]
        if abs(float(axv[2])) < 0.9:
            v = (axv[1], -axv[0], f32(0.0))
        else:
            v = (f32(0.0), axv[2], -axv[1])
        ln = np.sqrt(f32((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]))
        e1 = (f32(v[0] / ln), f32(v[1] / ln), f32(v[2] / ln))
        e2 = (f32(axv[1] * e1[2] - axv[2] * e1[1]), f32(axv[2] * e1[0] - axv[0] * e1[2]),
              f32(axv[0] * e1[1] - axv[1] * e1[0]))
        d2 = (ppx * e2[0] + ppy * e2[1]) + ppz * e2[2]
        d1 = (ppx * e1[0] + ppy * e1[1]) + ppz * e1[2]
        pi, tpi = f32(math.pi), f32(2 * math.pi)
        theta = np.arctan2(d2, d1).astype(np.float32) - tab["rot"]
        theta = np.where(theta < -pi, theta + tpi, theta)
        theta = np.where(theta > pi, theta - tpi, theta)
        return np.floor((theta + pi) / f32(2.094395102393195)).astype(np.int64) % 3

    if s_type == 2:
        return (dotA < 0) if sidx == 0 else (dotA >= 0)
    if s_type == 3:
        return sector() == sidx
    if s_type == 4:
        big = r_m > f32(1e-6)
        rr = np.where(big, r_m, f32(1.0))
        ur = (rx / rr, ry / rr, rz / rr)
        best_v = np.zeros(r_m.shape, np.int64)
        best_d = (ur[0] + ur[1]) + ur[2]
        for iv, (a, b, c) in enumerate(((1, -1, -1), (-1, 1, -1), (-1, -1, 1)), start=1):
            d = (f32(a) * ur[0] + f32(b) * ur[1]) + f32(c) * ur[2]
            better = d > best_d
            best_d = np.where(better, d, best_d)
            best_v = np.where(better, iv, best_v)
        return np.where(big, best_v == sidx, True)
    if s_type == 5:
        if sidx == 0:
            return dotA >= 0
        return (dotA < 0) & (sector() == sidx - 1)
    return np.ones(r_m.shape, bool)


def update(mf: MolFast, K, U, W, P):
    """The fused update (mol_fast.js:134-351) on the interior; returns new
    (U, W, P), faces copied from the inputs (the ping-pong partners hold the
    same faces)."""
    c = mf.consts()
    tab = mf.orbital_tables()
    atoms = mf.atoms_f32()
    S, NE = mf.S, mf.NE
    h, inv_h, inv_h2, dt = c["h"], c["inv_h"], c["inv_h2"], c["dt"]
    Ii, Jj, Kk = _coords(S)
    x_au = (Ii - f32(c["N2"])) * h
    y_au = (Jj - f32(c["N2"])) * h
    Kc = K[_interior(S)]
    un = [U[n][_interior(S)] for n in range(NE)]
    Pn = [P[n][_interior(S)] for n in range(NE)]
    P_total = np.zeros_like(Kc)
    for n in range(NE):
        P_total = P_total + Pn[n]
    Uo, Wo, Po = U.copy(), W.copy(), P.copy()
    for m in range(NE):
        uc = un[m]
        umax = np.zeros_like(uc)
        for n in range(NE):
            if n != m:
                umax = np.where(un[n] > umax, un[n], umax)
        cm = uc - umax
        w = W[m]
        wc = w[_interior(S)]
        wip, wim, wjp, wjm, wkp, wkm = (_nb(w, 1, 0, 0), _nb(w, -1, 0, 0), _nb(w, 0, 1, 0), _nb(w, 0, -1, 0),
                                        _nb(w, 0, 0, 1), _nb(w, 0, 0, -1))
        lw = (((((wip + wim) + wjp) + wjm) + wkp) + wkm - f32(6.0) * wc) * inv_h2
        gx, gy, gz = (wip - wim) * inv_h, (wjp - wjm) * inv_h, (wkp - wkm) * inv_h
        nw = (wc + ((f32(0.05) * dt) * np.abs(cm)) * lw) + ((f32(1.0) * dt) * cm) * np.sqrt(
            (gx * gx + gy * gy) + gz * gz)
        nw = np.minimum(np.maximum(nw, f32(0.0)), f32(1.0))

        def smooth_rc(nw, r, rc):
            edge = rc - f32(3.0) * h
            t = np.minimum(np.maximum((r - edge) / (rc - edge), f32(0.0)), f32(1.0))
            return np.where(r < rc, np.minimum(nw, (t * t) * (f32(3.0) - f32(2.0) * t)), nw)

        if mf.shell_mode:
            ax, ay, az, _ = atoms[m]
            rel = (Ii - ax, Jj - ay, Kk - az)
            r_m = np.sqrt((rel[0] * rel[0] + rel[1] * rel[1]) + rel[2] * rel[2]) * h
            rc_m = tab[m]["rc"]
            ro_m = tab[m]["r_out"] * c["R_out"]
            if rc_m > 0:
                nw = smooth_rc(nw, r_m, rc_m)
            if ro_m > 0:
                nw = np.where(r_m > ro_m, np.where(r_m <= ro_m + h, nw, f32(0.0)), nw)
            if tab[m]["split"] != 0 and tab[m]["fix"]:
                nw = np.where(_sector_mask_f32(m, tab[m], rel, r_m), nw, f32(0.0))
        else:
            for a in range(NE):
                ax, ay, az, _ = atoms[a]
                dx, dy, dz = Ii - ax, Jj - ay, Kk - az
                r = np.sqrt((dx * dx + dy * dy) + dz * dz) * h
                if tab[a]["rc"] > 0:
                    nw = smooth_rc(nw, r, tab[a]["rc"])
        Wo[m][_interior(S)] = nw

        P_group = np.zeros_like(uc)
        for n in range(NE):
            if tab[n]["group"] == tab[m]["group"]:
                P_group = P_group + Pn[n]
        selfN = tab[m]["target"]
        fm_multi = f32(0.0) if c["full_self"] > 0.5 else f32(f32(1.0) / selfN)
        fm = fm_multi if selfN > 1.0 else f32(1.0)
        Vother = f32(2.0) * (P_total - fm * P_group)
        rho_m = uc * uc
        if selfN > 1.0 and c["alpha_tf"] > 0.0:
            V_TF = c["alpha_tf"] * np.power(np.maximum(rho_m, f32(1e-20)), f32(0.6666667))
        else:
            V_TF = np.zeros_like(uc)
        V_tilt = (-tab[m]["tilt_x"]) * x_au - tab[m]["tilt_y"] * y_au
        rho_other = np.zeros_like(uc)
        for n in range(NE):
            if n != m:
                rho_other = rho_other + un[n] * un[n]
        V_pauli = c["beta"] * rho_other
        u = U[m]
        uip, uim, ujp, ujm, ukp, ukm = (_nb(u, 1, 0, 0), _nb(u, -1, 0, 0), _nb(u, 0, 1, 0), _nb(u, 0, -1, 0),
                                        _nb(u, 0, 0, 1), _nb(u, 0, 0, -1))
        k_mult = c["kc_multi"] if selfN > 1.0 else f32(1.0)
        hd = c["half_d"] * k_mult
        half = f32(0.5)
        A = hd * (((uip - uc) * (wip + nw)) * half - ((uc - uim) * (nw + wim)) * half)
        B = hd * (((ujp - uc) * (wjp + nw)) * half - ((uc - ujm) * (nw + wjm)) * half)
        C = hd * (((ukp - uc) * (wkp + nw)) * half - ((uc - ukm) * (nw + wkm)) * half)
        D = ((dt * (((((Kc - Vother) - V_TF) - V_tilt) - V_pauli) - c["field_x"] * x_au)) * uc) * wc
        Uo[m][_interior(S)] = (((uc + A) + B) + C) + D
        p = P[m]
        Pc = Pn[m]
        s6 = (((((_nb(p, 1, 0, 0) + _nb(p, -1, 0, 0)) + _nb(p, 0, 1, 0)) + _nb(p, 0, -1, 0)) + _nb(p, 0, 0, 1))
              + _nb(p, 0, 0, -1))
        Po[m][_interior(S)] = (Pc + (dt * (s6 - f32(6.0) * Pc)) * inv_h2) + ((c["two_pi"] * dt) * uc) * uc
    return Uo, Wo, Po


def reduce_sums(mf: MolFast, K, U, W, P):
    """reduce + finalize (mol_fast.js:352-466): [norm_0..norm_{NE-1}, T, V_eK,
    V_ee, dipEx, dipEy, dipEz], per-node FP32 terms summed in float64."""
    c = mf.consts()
    tab = mf.orbital_tables()
    S, NE = mf.S, mf.NE
    h, h3 = c["h"], c["h3"]
    Ii, Jj, Kk = _coords(S)
    xc, yc, zc = (Ii - f32(c["N2"])) * h, (Jj - f32(c["N2"])) * h, (Kk - f32(c["N2"])) * h
    Kc = K[_interior(S)]
    P_tot = np.zeros_like(Kc)
    for n in range(NE):
        P_tot = P_tot + P[n][_interior(S)]
    T = np.zeros_like(Kc)
    VeK = np.zeros_like(Kc)
    Vee = np.zeros_like(Kc)
    dip = [np.zeros_like(Kc) for _ in range(3)]
    norms = []
    for m in range(NE):
        u = U[m]
        v = u[_interior(S)]
        norms.append(float(np.sum(((v * v) * h3).astype(np.float64))))
        selfN = tab[m]["target"]
        k_mult = c["kc_multi"] if selfN > 1.0 else f32(1.0)
        a, b, cc = _nb(u, 1, 0, 0) - v, _nb(u, 0, 1, 0) - v, _nb(u, 0, 0, 1) - v
        tterm = ((k_mult * c["kc"]) * ((a * a + b * b) + cc * cc)) * h
        T = np.where(W[m][_interior(S)] > f32(0.1), T + tterm, T)
        VeK = VeK - ((Kc * v) * v) * h3
        rho_cell = (v * v) * h3
        dip[0] = dip[0] - rho_cell * xc
        dip[1] = dip[1] - rho_cell * yc
        dip[2] = dip[2] - rho_cell * zc
        fm = f32(f32(1.0) / selfN) if selfN > 1.0 else f32(1.0)
        Vee = Vee + (((P_tot - fm * P[m][_interior(S)]) * v) * v) * h3
        if selfN > 1.0 and c["alpha_tf"] > 0.0:
            T = T + ((f32(0.6) * c["alpha_tf"]) * np.power(np.maximum(v * v, f32(1e-20)), f32(1.6666667))) * h3
        tilt = tab[m]["tilt_y"]
        if tilt != 0.0:
            T = T + ((((-tilt) * yc) * v) * v) * h3
    rest = [float(np.sum(x.astype(np.float64))) for x in (T, VeK, Vee, dip[0], dip[1], dip[2])]
    return np.array(norms + rest)


def normalize(mf: MolFast, U, sums):
    """normalize (mol_fast.js:470-489): U_m *= sqrt(target / n_m) on the interior."""
    tab = mf.orbital_tables()
    out = U.copy()
    for m in range(mf.NE):
        n = f32(sums[m])
        tgt = tab[m]["target"]
        if n > 0 and tgt > 0:
            out[m][_interior(mf.S)] = U[m][_interior(mf.S)] * np.sqrt(f32(tgt / n))
    return out


def restrict_fw(mf: MolFast, fine):
    """Full weighting fine -> coarse interior (mol_fast.js:690-720)."""
    S = mf.S
    NC = mf.NN // 2
    SC = NC + 1
    co = np.zeros((SC, SC, SC), np.float32)
    f = fine

    def at(di, dj, dk):
        return f[2 + di:2 * NC - 1 + di:2, 2 + dj:2 * NC - 1 + dj:2, 2 + dk:2 * NC - 1 + dk:2]

    val = f32(8.0) * at(0, 0, 0)
    faces = (((((at(1, 0, 0) + at(-1, 0, 0)) + at(0, 1, 0)) + at(0, -1, 0)) + at(0, 0, 1)) + at(0, 0, -1))
    val = val + f32(4.0) * faces
    edges = [at(1, 1, 0), at(1, -1, 0), at(-1, 1, 0), at(-1, -1, 0), at(1, 0, 1), at(1, 0, -1),
             at(-1, 0, 1), at(-1, 0, -1), at(0, 1, 1), at(0, 1, -1), at(0, -1, 1), at(0, -1, -1)]
    e = edges[0]
    for x in edges[1:]:
        e = e + x
    val = val + f32(2.0) * e
    corners = [at(1, 1, 1), at(1, 1, -1), at(1, -1, 1), at(1, -1, -1), at(-1, 1, 1), at(-1, 1, -1),
               at(-1, -1, 1), at(-1, -1, -1)]
    cs = corners[0]
    for x in corners[1:]:
        cs = cs + x
    val = val + cs
    co[1:NC, 1:NC, 1:NC] = val * f32(0.015625)
    del S
    return co


def coarse_smooth(mf: MolFast, E, rhs):
    """One coarse weighted-Jacobi sweep (mol_fast.js:722-745)."""
    c = mf.consts()
    out = E.copy()
    ec = E[1:-1, 1:-1, 1:-1]
    s = (((((_nb(E, 1, 0, 0) + _nb(E, -1, 0, 0)) + _nb(E, 0, 1, 0)) + _nb(E, 0, -1, 0)) + _nb(E, 0, 0, 1))
         + _nb(E, 0, 0, -1))
    hc2 = f32(4.0) * c["h2"]
    out[1:-1, 1:-1, 1:-1] = f32(0.3333) * ec + (s + hc2 * rhs[1:-1, 1:-1, 1:-1]) / f32(9.0)
    return out


def prolong_correct(mf: MolFast, Pm, Ec):
    """P_m += 0.5 trilinear(Ec) on the fine interior (mol_fast.js:747-789)."""
    S, NC = mf.S, mf.NN // 2
    SC = NC + 1
    i = np.arange(1, S - 1)
    ci, ci1 = i // 2, np.minimum(i // 2 + (i & 1), NC)
    w = np.where(i & 1, f32(0.5), f32(1.0)).astype(np.float32)
    w1 = (f32(1.0) - w).astype(np.float32)
    E = Ec.reshape(SC, SC, SC)
    I_, J_, K_ = np.ix_(range(len(i)), range(len(i)), range(len(i)))
    wi, wj, wk = w[I_], w[J_], w[K_]
    wi1, wj1, wk1 = w1[I_], w1[J_], w1[K_]
    a0, a1 = ci[I_], ci1[I_]
    b0, b1 = ci[J_], ci1[J_]
    c0, c1 = ci[K_], ci1[K_]
    corr = np.zeros((len(i),) * 3, np.float32)
    for (wa, ia), (wb, ib), (wc, ic) in (((wi, a0), (wj, b0), (wk, c0)), ((wi, a0), (wj, b0), (wk1, c1)),
                                         ((wi, a0), (wj1, b1), (wk, c0)), ((wi, a0), (wj1, b1), (wk1, c1)),
                                         ((wi1, a1), (wj, b0), (wk, c0)), ((wi1, a1), (wj, b0), (wk1, c1)),
                                         ((wi1, a1), (wj1, b1), (wk, c0)), ((wi1, a1), (wj1, b1), (wk1, c1))):
        corr = corr + ((wa * wb) * wc) * E[ia, ib, ic]
    out = Pm.copy()
    out[1:-1, 1:-1, 1:-1] = Pm[1:-1, 1:-1, 1:-1] + f32(0.5) * corr
    return out


def twogrid(mf: MolFast, U, P):
    """Per-electron 2-grid correction of every P_m (mol_fast.js:1262-1287)."""
    c = mf.consts()
    S, NC = mf.S, mf.NN // 2
    SC = NC + 1
    out = P.copy()
    for m in range(mf.NE):
        u, p = U[m], P[m]
        rho = u * u
        Pc = p[_interior(S)]
        s6 = (((((_nb(p, 1, 0, 0) + _nb(p, -1, 0, 0)) + _nb(p, 0, 1, 0)) + _nb(p, 0, -1, 0)) + _nb(p, 0, 0, 1))
              + _nb(p, 0, 0, -1))
        lap = (s6 - f32(6.0) * Pc) * c["inv_h2"]
        res = np.zeros((S, S, S), np.float32)
        res[_interior(S)] = (-c["two_pi"]) * rho[_interior(S)] - lap
        rhs = restrict_fw(mf, res)
        e = [np.zeros((SC, SC, SC), np.float32), np.zeros((SC, SC, SC), np.float32)]
        for cs in range(mf.coarse_iters):
            src, dst = (0, 1) if cs % 2 == 0 else (1, 0)
            e[dst] = coarse_smooth(mf, e[src], rhs)
        out[m] = prolong_correct(mf, p, e[0])
    return out


def do_steps(mf: MolFast, K, U, W, P, n, t_step=0):
    """doSteps(n) (mol_fast.js:1237-1322): n fused updates with normalisation
    every norm_interval steps and at the last step, 2-grid corrections every
    vcycle_interval steps.  Returns (U, W, P, sums of the last reduction)."""
    sums = None
    for s in range(n):
        U, W, P = update(mf, K, U, W, P)
        if (s + 1) % mf.norm_interval == 0 or s == n - 1:
            sums = reduce_sums(mf, K, U, W, P)
            U = normalize(mf, U, sums)
        if mf.vcycle and (s + 1) % mf.vcycle_interval == 0:
            P = twogrid(mf, U, P)
    return U, W, P, sums


def forces_elec(mf: MolFast, P, atoms=None):
    """Per-atom sphere-averaged gradient of sum_m P_m, F = 2 Z <grad P>
    (mol_fast.js:496-560), FP32 in the shader's loop order."""
    c = mf.consts()
    S = mf.S
    atoms = atoms if atoms is not None else mf.atoms_f32()
    inv2h = f32(0.5) * c["inv_h"]
    R = FORCE_RADIUS
    out = np.zeros((len(atoms), 3), np.float32)
    Psum_cache = {}
    for a, (ax, ay, az, Z) in enumerate(atoms):
        if not Z > 0:
            continue
        ci, cj, ck = int(np.trunc(ax)), int(np.trunc(ay)), int(np.trunc(az))
        sF = [f32(0.0)] * 3
        count = f32(0.0)
        for di in range(-R, R + 1):
            ii = ci + di
            if ii < 1 or ii >= S - 1:
                continue
            for dj in range(-R, R + 1):
                jj = cj + dj
                if jj < 1 or jj >= S - 1:
                    continue
                for dk in range(-R, R + 1):
                    kk = ck + dk
                    if kk < 1 or kk >= S - 1:
                        continue
                    if di * di + dj * dj + dk * dk > R * R:
                        continue
                    vals = []
                    for (oi, oj, ok) in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                        key = (ii + oi, jj + oj, kk + ok)
                        if key not in Psum_cache:
                            acc = f32(0.0)
                            for m in range(P.shape[0]):
                                acc = f32(acc + P[m][key])
                            Psum_cache[key] = acc
                        vals.append(Psum_cache[key])
                    sF[0] = f32(sF[0] + f32(f32(vals[0] - vals[1]) * inv2h))
                    sF[1] = f32(sF[1] + f32(f32(vals[2] - vals[3]) * inv2h))
                    sF[2] = f32(sF[2] + f32(f32(vals[4] - vals[5]) * inv2h))
                    count = f32(count + f32(1.0))
        avg = f32(f32(1.0) / count) if count > 0 else f32(0.0)
        for d in range(3):
            out[a, d] = f32(f32(f32(f32(2.0) * Z) * sF[d]) * avg)
    return out


def recompute_k(mf: MolFast, atoms=None):
    """recomputeK (mol_fast.js:563-586), FP32 on all nodes, softening 4 h^2."""
    c = mf.consts()
    S = mf.S
    atoms = atoms if atoms is not None else mf.atoms_f32()
    ax = np.arange(S, dtype=np.float32)
    I, J, Kk = np.meshgrid(ax, ax, ax, indexing="ij")
    K = np.zeros((S, S, S), np.float32)
    soft = f32(4.0) * c["h2"]
    for (x, y, z, Z) in atoms:
        if not Z > 0:
            continue
        di, dj, dk = (I - x) * c["h"], (J - y) * c["h"], (Kk - z) * c["h"]
        K = K + Z / np.sqrt(((di * di + dj * dj) + dk * dk) + soft)
    return K


def extract(mf: MolFast, U, W, P):
    """Slice extract (mol_fast.js:592-641): 3 NELEC S^2 floats."""
    c = mf.consts()
    S, NE, N2 = mf.S, mf.NE, c["N2"]
    tab = mf.orbital_tables()
    out = np.zeros(3 * NE * S * S, np.float32)
    for m in range(NE):
        u2 = U[m][:, :, N2]
        w2 = W[m][:, :, N2]
        out[m * S * S:(m + 1) * S * S] = np.where(w2 > f32(0.1), u2, f32(0.0)).ravel()
        out[NE * S * S + m * S:NE * S * S + (m + 1) * S] = U[m][:, N2, N2]
        out[2 * NE * S * S + m * S:2 * NE * S * S + (m + 1) * S] = W[m][:, N2, N2]
        if m == 0:
            psum = np.zeros(S, np.float32)
            for n in range(NE):
                psum = psum + P[n][:, N2, N2]
            out[NE * S * S + NE * S:NE * S * S + NE * S + S] = f32(2.0) * psum
        vp = np.zeros(S, np.float32)
        if tab[m]["target"] > 1.0 and c["alpha_tf"] > 0.0:
            rho = U[m][:, N2, N2] * U[m][:, N2, N2]
            vp = c["alpha_tf"] * np.power(np.maximum(rho, f32(1e-20)), f32(0.6666667))
        base = NE * S * S + (NE + 1 + m) * S
        out[base:base + S] = vp
    return out


# ---------------------------------------------------------------------------
# host-side nuclear terms (JavaScript doubles)
# ---------------------------------------------------------------------------
def e_kk(mf: MolFast, pos=None):
    """computeEKK (mol_fast.js:1326-1337)."""
    hv = mf.screen / mf.NN
    pos = pos if pos is not None else [(n.i, n.j, n.k) for n in mf.nuclei]
    E = 0.0
    for a in range(len(pos)):
        for b in range(a + 1, len(pos)):
            dx, dy, dz = ((pos[a][d] - pos[b][d]) * hv for d in range(3))
            r = math.sqrt(dx * dx + dy * dy + dz * dz)
            if r > 0:
                E += mf.nuclei[a].Z * mf.nuclei[b].Z / r
    return E


def forces_total(mf: MolFast, felec, pos=None):
    """computeForces (mol_fast.js:1339-1373): electronic part plus the
    softened (epsilon^2 = h^2) nucleus-nucleus force."""
    hv = mf.screen / mf.NN
    h2v = hv * hv
    pos = pos if pos is not None else [(n.i, n.j, n.k) for n in mf.nuclei]
    out = np.array(felec, np.float64).reshape(-1, 3).copy()
    for a in range(len(pos)):
        for b in range(len(pos)):
            if b == a:
                continue
            dx, dy, dz = ((pos[a][d] - pos[b][d]) * hv for d in range(3))
            r2 = dx * dx + dy * dy + dz * dz + h2v
            pref = mf.nuclei[a].Z * mf.nuclei[b].Z * (1 / (r2 * math.sqrt(r2)))
            out[a] += (pref * dx, pref * dy, pref * dz)
    return out
