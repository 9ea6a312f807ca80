"""S1: CPU restatement of the ``cloud/solver.py`` engine (TEST INFRASTRUCTURE ONLY).

Grid convention (``cloud/solver.py:322-329``): ``NN`` cells and ``S = NN + 1``
nodes per axis, spacing ``h = screen / NN``; arrays here are ``(S, S, S)``
indexed ``[i, j, k]`` which is the reference's flat ``id = i*S*S + j*S + k``.
Face nodes (any index 0 or NN) are Dirichlet data: every function below writes
interior nodes only and carries faces through unchanged.  In the S1 engine
the faces of U, P, K and label are all zero (``solver.py:27-29,73-75,201-203``),
so this is the same as the reference re-zeroing them.

Atoms are ``(n, 5)`` rows ``(i, j, k, Z, rc)`` in fractional grid units
(``solver.py:378-382``).

dtype ``np.float64`` = FP64 oracle; ``np.float32`` = FP32-faithful mode that
performs the reference kernel's single-precision operations in the same
order (the HIP kernels are compiled with ``-ffp-contract=off`` to match).
Reductions (norms, energies, forces) accumulate in FP64 in both modes -- the
reference uses FP32 ``scatter_add`` / ``atomicAdd`` there (documented
deviation: order-nondeterministic FP32 sums replaced by FP64 sums).

Parity pinning (tests/test_oracle_s1.py):
  init_trial     closed form per voxel by scalar brute force; argmax first-index tie rule
  itp_update     dense-matrix assembly of the (linear for fixed K,P,label) operator on 8^3
  jacobi         fixed point = scipy sparse direct solve of the 7-point Dirichlet system
  normalize      invariant sum_{label=d} U^2 h^3 = Z_d after normalise
  boundary_score brute-force scalar rule on a tiny grid; flips only at foreign-neighbour voxels
  energy         brute-force scalar loops; sum of per-domain parts = totals
  hf_forces      brute-force scalar sum; translation/Newton checks on the nuclear part
  move_nuclei    scalar closed form
  run            hydrogen single domain -> E -> -0.5 (coarse tolerance, h -> 0 trend)
"""
from __future__ import annotations

import math
import numpy as np

INV_SQRT_PI = 0.5641895835  # solver.py:35 literal
TWO_PI = 2.0 * math.pi      # solver.py:437


def dv_table(n_elec: int) -> float:
    """ITP step dv by active-electron count (solver.py:341-351)."""
    if n_elec > 500:
        return 0.005
    if n_elec > 200:
        return 0.01
    if n_elec > 100:
        return 0.02
    if n_elec > 30:
        return 0.03
    return 0.12


class Grid:
    """Scalar grid constants as the reference hands them to its kernels.

    ``solver.py`` computes h, h^2, h^3, dv*h^2 in Python double precision and
    casts each to float32 at the kernel call (``np.float32(self.h2)`` etc.), so
    the FP32 constants are ``float32(h*h)``, not ``float32(h)**2``.
    """

    def __init__(self, NN: int, screen: float, dv: float, dtype=np.float64):
        self.NN = int(NN)
        self.S = self.NN + 1
        self.dtype = dtype
        h = screen / NN
        self.h64 = h
        self.h = dtype(h)
        self.h2 = dtype(h * h)
        self.h3 = dtype(h ** 3)
        self.dv = dv
        self.half_dv = dtype(0.5 * dv)
        self.dtv = dtype(dv * (h * h))
        self.two_pi = dtype(TWO_PI)


def _interior(S):
    return (slice(1, S - 1),) * 3


def _nb(a, axis, sign):
    """Neighbour view of the interior block of ``a`` shifted by +-1 along axis."""
    S = a.shape[0]
    sl = [slice(1, S - 1)] * 3
    sl[axis] = slice(2, S) if sign > 0 else slice(0, S - 2)
    return a[tuple(sl)]


def init_trial(atoms, g: Grid):
    """Trial state: K, Voronoi-by-trial label, U = best trial (solver.py:16-59).

    K = sum_n Z_n / r~_n; trial u_n = Z_n^2/sqrt(pi) exp(-Z_n r~_n); label = first
    argmax (strict ``>`` from bU = 0, bestN = 0); U = bestU = max trial.
    r~ = max(sqrt(r^2 + 0.04 h^2), rc) if rc > 0 else sqrt(r^2 + h^2).
    Returns (U, label[int32], K, bestU); faces zero.
    """
    dt = g.dtype
    S = g.S
    n = S - 2
    atoms = np.asarray(atoms, dtype=dt).reshape(-1, 5)
    x = np.arange(1, S - 1).astype(dt) * g.h
    bU = np.zeros((n, n, n), dt)
    best = np.zeros((n, n, n), np.int32)
    Kv = np.zeros((n, n, n), dt)
    c004 = dt(0.04)
    inv_sqrt_pi = dt(INV_SQRT_PI)
    for a in range(atoms.shape[0]):
        ai, aj, ak, Za, rc = atoms[a]
        if not Za > 0:
            continue
        dx = (x - ai * g.h)[:, None, None]
        dy = (x - aj * g.h)[None, :, None]
        dz = (x - ak * g.h)[None, None, :]
        r2 = dx * dx + dy * dy + dz * dz
        if rc > 0:
            r = np.maximum(np.sqrt(r2 + c004 * g.h2), rc)
        else:
            r = np.sqrt(r2 + g.h2)
        Kv = Kv + Za / r
        ut = Za * Za * inv_sqrt_pi * np.exp(-Za * r)
        upd = ut > bU
        bU = np.where(upd, ut, bU)
        best = np.where(upd, np.int32(a), best)
    U = np.zeros((S, S, S), dt)
    K = np.zeros((S, S, S), dt)
    label = np.zeros((S, S, S), np.int32)
    I = _interior(S)
    U[I] = bU
    K[I] = Kv
    label[I] = best
    return U, label, K, U.copy()


def recompute_k(atoms, g: Grid):
    """K refresh after nuclear motion (solver.py:149-179): same r~ as init."""
    dt = g.dtype
    S = g.S
    n = S - 2
    atoms = np.asarray(atoms, dtype=dt).reshape(-1, 5)
    x = np.arange(1, S - 1).astype(dt) * g.h
    Kv = np.zeros((n, n, n), dt)
    c004 = dt(0.04)
    for a in range(atoms.shape[0]):
        ai, aj, ak, Za, rc = atoms[a]
        if not Za > 0:
            continue
        dx = (x - ai * g.h)[:, None, None]
        dy = (x - aj * g.h)[None, :, None]
        dz = (x - ak * g.h)[None, None, :]
        r2 = dx * dx + dy * dy + dz * dz
        if rc > 0:
            r = np.maximum(np.sqrt(r2 + c004 * g.h2), rc)
        else:
            r = np.sqrt(r2 + g.h2)
        Kv = Kv + Za / r
    K = np.zeros((S, S, S), dt)
    K[_interior(S)] = Kv
    return K


def itp_update(U, K, P, label, g: Grid, rc_mask=None):
    """One label-Neumann imaginary-time step (solver.py:61-90).

    U' = U + (dv/2) * sum_6 (u~_q - U) + dv h^2 (K - 2P) U with u~_q = U_q when
    label_q == label else U (Neumann mirror at domain walls).  The Laplacian is
    summed in the reference order (+i, -i, +j, -j, +k, -k) then ``- 6U``.

    ``rc_mask`` (optional, bool S^3) adds the molecule.js pseudopotential rule
    (molecule.js:228-343): U' = 0 where the own mask is set, and a masked
    neighbour is treated like a foreign one.
    """
    dt = g.dtype
    S = g.S
    I = _interior(S)
    c = U[I]
    L = label[I]
    if rc_mask is not None:
        mc = rc_mask[I]
    terms = []
    for axis in range(3):
        for sign in (1, -1):
            ln = _nb(label, axis, sign)
            same = ln == L
            if rc_mask is not None:
                same = same & ~_nb(rc_mask, axis, sign)
            terms.append(np.where(same, _nb(U, axis, sign), c))
    lap = terms[0] + terms[1] + terms[2] + terms[3] + terms[4] + terms[5] - dt(6.0) * c
    new = c + g.half_dv * lap + g.dtv * (K[I] - dt(2.0) * P[I]) * c
    if rc_mask is not None:
        new = np.where(mc, dt(0.0), new)
    out = U.copy()
    out[I] = new
    return out


def sum6(P):
    """Six-neighbour sum in the reference order +i, -i, +j, -j, +k, -k."""
    return (_nb(P, 0, 1) + _nb(P, 0, -1) + _nb(P, 1, 1) + _nb(P, 1, -1)
            + _nb(P, 2, 1) + _nb(P, 2, -1))


def smooth(P, rhs, a, c, dtype):
    """Generic damped-Jacobi sweep P' = a*P + (sum_6 P + rhs) / c (interior).

    a = 0, c = 6 is the S1 Jacobi step (solver.py:190-208); a = 0.3333, c = 9
    is the molecule.js smoother (molecule.js:723-749).  ``rhs`` is the interior
    right-hand side already scaled by h^2.
    """
    dt = dtype
    S = P.shape[0]
    I = _interior(S)
    s = sum6(P) + rhs
    new = s / dt(c)
    if a != 0:
        new = dt(a) * P[I] + new
    out = P.copy()
    out[I] = new
    return out


def rho_rhs(U, alpha, beta, dtype):
    """Fine-level Poisson right-hand side (alpha * U^2) * beta on the interior.

    S1 (solver.py:206): alpha = 2*pi, beta = h^2, i.e. ``coeff * rho * h2``
    with rho = U*U stored in FP32 (solver.py:186).
    """
    I = _interior(U.shape[0])
    u = U[I]
    return (dtype(alpha) * (u * u)) * dtype(beta)


def jacobi(P, U, g: Grid, sweeps=2):
    """S1 Poisson step: rho = U^2, then ``sweeps`` Jacobi sweeps (solver.py:434-444)."""
    rhs = rho_rhs(U, g.two_pi, g.h2, g.dtype)
    for _ in range(sweeps):
        P = smooth(P, rhs, 0.0, 6.0, g.dtype)
    return P


def domain_norms(U, label, n_atoms, g: Grid):
    """n_d = sum over interior nodes with clip(label, 0, n-1) == d of U*U*h^3.

    cloud/solver.py:421-427 sums all nodes, but its face values of U are
    always zero, so the interior sum is the same; molecule.js sums interior
    nodes only (molecule.js:912-935), where the faces hold trial values.
    The per-node term is formed in the working dtype; the sum is FP64.
    """
    I = _interior(U.shape[0])
    t = (U[I] * U[I]) * g.h3
    lab = np.clip(label[I], 0, n_atoms - 1).ravel()
    return np.bincount(lab, weights=t.ravel().astype(np.float64), minlength=n_atoms)


def norm_scales(norms, Z, dtype):
    """Per-domain scale sqrt(Z_d / n_d); 1 where Z_d <= 0 or n_d <= 1e-20 (solver.py:210-225)."""
    Z = np.asarray(Z, np.float64)
    sc = np.ones(len(norms), np.float64)
    ok = (Z > 0) & (norms > 1e-20)
    sc[ok] = np.sqrt(Z[ok] / norms[ok])
    return sc.astype(dtype)


def normalize(U, label, Z, g: Grid):
    """Per-domain renormalisation U *= sqrt(Z_d / n_d) on interior nodes
    (solver.py:421-432; molecule.js:1077-1105 for the interior-only rule)."""
    n = len(Z)
    norms = domain_norms(U, label, n, g)
    sc = norm_scales(norms, Z, g.dtype)
    I = _interior(U.shape[0])
    L = label[I]
    ok = (L >= 0) & (L < n)
    out = U.copy()
    out[I] = np.where(ok, U[I] * sc[np.clip(L, 0, n - 1)], U[I])
    return out, norms


def boundary_score(U, label, atoms, g: Grid, ratio=1.1, min_foreign=2):
    """Nuclear-score boundary flip (solver.py:92-147), double-buffered.

    At interior voxels with >= 1 foreign neighbour: bestN = first argmax over
    active atoms of Z/sqrt(d^2 + h^2) (strict ``>``, from score 0 / current
    label); flip when bestN != label, best > ratio * score(label) and the
    foreign-neighbour count >= ``min_foreign``.  All neighbour reads come from
    the input labels (the reference updates in place, which races; the
    double-buffered rule is its deterministic reading).
    """
    dt = g.dtype
    S = g.S
    I = _interior(S)
    L = label[I]
    other = np.zeros(L.shape, np.int32)
    for axis in range(3):
        for sign in (1, -1):
            other += (_nb(label, axis, sign) != L).astype(np.int32)
    out = label.copy()
    cand = np.argwhere(other > 0)
    if cand.size == 0:
        return out
    atoms = np.asarray(atoms, dtype=dt).reshape(-1, 5)
    ci = cand[:, 0] + 1
    cj = cand[:, 1] + 1
    ck = cand[:, 2] + 1
    xi = ci.astype(dt) * g.h
    yj = cj.astype(dt) * g.h
    zk = ck.astype(dt) * g.h
    cur = L[cand[:, 0], cand[:, 1], cand[:, 2]]
    cur_score = np.zeros(len(cand), dt)
    best_score = np.zeros(len(cand), dt)
    best_n = cur.copy()
    for a in range(atoms.shape[0]):
        ai, aj, ak, Za, _rc = atoms[a]
        if not Za > 0:
            continue
        dx = xi - ai * g.h
        dy = yj - aj * g.h
        dz = zk - ak * g.h
        r2 = dx * dx + dy * dy + dz * dz + g.h2
        sc = Za / np.sqrt(r2)
        cur_score = np.where(cur == a, sc, cur_score)
        upd = sc > best_score
        best_score = np.where(upd, sc, best_score)
        best_n = np.where(upd, a, best_n)
    oc = other[cand[:, 0], cand[:, 1], cand[:, 2]]
    flip = (best_n != cur) & (best_score > cur_score * dt(ratio)) & (oc >= min_foreign)
    out[ci[flip], cj[flip], ck[flip]] = best_n[flip]
    return out


def energy(U, K, P, label, g: Grid, n_domains=None):
    """T, V_eK, V_ee totals and per-domain parts (solver.py:227-280).

    T   = sum 0.5 * (du_i^2 + du_j^2 + du_k^2) * h with forward differences
          counted only toward same-label neighbours;
    VeK = sum -K * rho * h^3;  Vee = sum P * rho * h^3;  rho = U^2.
    Per-voxel terms in the working dtype, sums in FP64.  Per-domain parts are
    keyed by the voxel's own label.  Returns (totals[3], per_domain[n, 3]).
    """
    dt = g.dtype
    S = g.S
    I = _interior(S)
    u = U[I]
    L = label[I]
    rho = u * u
    d = []
    for axis in range(3):
        d.append(np.where(_nb(label, axis, 1) == L, _nb(U, axis, 1) - u, dt(0.0)))
    T = dt(0.5) * (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) * g.h
    VeK = -K[I] * rho * g.h3
    Vee = P[I] * rho * g.h3
    if n_domains is None:
        n_domains = int(label.max()) + 1
    lab = L.ravel()
    per = np.stack([np.bincount(lab, weights=t.ravel().astype(np.float64), minlength=n_domains)
                    for t in (T, VeK, Vee)], axis=1)
    tot = np.array([T.astype(np.float64).sum(), VeK.astype(np.float64).sum(),
                    Vee.astype(np.float64).sum()])
    return tot, per


def nuclear_energy(pos, Z, h, soft=1.0):
    """E_KK = sum_{a<b} Z_a Z_b / sqrt(R_ab^2 + soft*h^2) over active nuclei (solver.py:462-474), FP64."""
    pos = np.asarray(pos, np.float64) * h
    Z = np.asarray(Z, np.float64)
    e = 0.0
    n = len(Z)
    for a in range(n):
        if Z[a] <= 0:
            continue
        for b in range(a + 1, n):
            if Z[b] <= 0:
                continue
            d = pos[a] - pos[b]
            e += Z[a] * Z[b] / math.sqrt(float(d @ d) + soft * h * h)
    return e


def nuclear_forces(pos, Z, h, n_force=None):
    """Nucleus-nucleus force on each of the first n_force atoms (solver.py:504-520), FP64."""
    pos = np.asarray(pos, np.float64) * h
    Z = np.asarray(Z, np.float64)
    n = len(Z)
    n_force = n if n_force is None else n_force
    F = np.zeros((n, 3))
    for a in range(n_force):
        if Z[a] <= 0:
            continue
        for b in range(n):
            if b == a or Z[b] <= 0:
                continue
            d = pos[a] - pos[b]
            r2 = float(d @ d) + h * h
            F[a] += Z[a] * Z[b] / (r2 * math.sqrt(r2)) * d
    return F


def hf_forces(U, atoms, g: Grid, first=0, count=None):
    """Electronic Hellmann-Feynman force (solver.py:282-315).

    F_n = Z_n sum_v rho_v h^3 d_v / (d_v^2 + h^2)^{3/2}, d = x_v - R_n, skipping
    voxels with rho < 1e-20; per-voxel terms in the working dtype, FP64 sums.
    """
    dt = g.dtype
    S = g.S
    I = _interior(S)
    atoms = np.asarray(atoms, dtype=dt).reshape(-1, 5)
    if count is None:
        count = atoms.shape[0] - first
    u = U[I]
    rho = u * u
    live = rho >= dt(1e-20)
    x = np.arange(1, S - 1).astype(dt) * g.h
    F = np.zeros((count, 3))
    for m in range(count):
        ai, aj, ak, Za, _rc = atoms[first + m]
        if not Za > 0:
            continue
        dx = np.broadcast_to((x - ai * g.h)[:, None, None], rho.shape)
        dy = np.broadcast_to((x - aj * g.h)[None, :, None], rho.shape)
        dz = np.broadcast_to((x - ak * g.h)[None, None, :], rho.shape)
        r2 = dx * dx + dy * dy + dz * dz + g.h2
        inv_r3 = dt(1.0) / (r2 * np.sqrt(r2))
        w = Za * rho * inv_r3 * g.h3
        for c, d in enumerate((dx, dy, dz)):
            F[m, c] = (w * d)[live].astype(np.float64).sum()
    return F


MASS = {1: 1 * 1836, 2: 16 * 1836, 3: 14 * 1836, 4: 12 * 1836}


def move_nuclei(pos, vel, force, Z, h, NN, dt_nuc, force_scale, damping, max_vel, n_move=None):
    """Damped explicit nuclear step (solver.py:522-537), FP64 restatement.

    v += F/m * dt * scale; v *= damping; v = clip(v, +-max_vel);
    x += v * dt / h; x = clip(x, 5, NN - 5) (grid units).  Masses by Z.
    """
    pos = np.array(pos, np.float64)
    vel = np.array(vel, np.float64)
    n = len(Z)
    n_move = n if n_move is None else n_move
    for a in range(n_move):
        z = int(Z[a])
        if z <= 0:
            continue
        m = MASS.get(z, 1836)
        for d in range(3):
            vel[a, d] += force[a][d] / m * dt_nuc * force_scale
            vel[a, d] *= damping
            vel[a, d] = min(max(vel[a, d], -max_vel), max_vel)
            pos[a, d] += vel[a, d] * dt_nuc / h
            pos[a, d] = min(max(pos[a, d], 5), NN - 5)
    return pos, vel


class State:
    """Mutable S1 solver state for the schedule restatement."""

    def __init__(self, atoms, NN, screen, dtype=np.float64, dv=None):
        atoms = np.asarray(atoms, np.float64).reshape(-1, 5)
        self.atoms = atoms
        self.Z = atoms[:, 3].copy()
        n_elec = int((self.Z > 0).sum())
        self.g = Grid(NN, screen, dv_table(n_elec) if dv is None else dv, dtype)
        S = self.g.S
        self.U = np.zeros((S, S, S), dtype)
        self.P = np.zeros((S, S, S), dtype)
        self.K = np.zeros((S, S, S), dtype)
        self.label = np.zeros((S, S, S), np.int32)

    def initialize(self, bootstrap=100, sweeps=2):
        """initTrial, normalise, ``bootstrap`` Poisson steps (solver.py:402-419)."""
        self.U, self.label, self.K, _ = init_trial(self.atoms, self.g)
        self.U, _ = normalize(self.U, self.label, self.Z, self.g)
        for _ in range(bootstrap):
            self.P = jacobi(self.P, self.U, self.g, sweeps)

    def step(self, step, poisson_interval=2, norm_interval=20, boundary=True, sweeps=2):
        """One iteration of the run() loop body without forces/report (solver.py:561-586)."""
        if step % poisson_interval == 0:
            self.P = jacobi(self.P, self.U, self.g, sweeps)
        self.U = itp_update(self.U, self.K, self.P, self.label, self.g)
        if step % norm_interval == 0:
            self.U, _ = normalize(self.U, self.label, self.Z, self.g)
        if boundary and step % (norm_interval * 5) == 0:
            self.label = boundary_score(self.U, self.label, self.atoms, self.g)

    def energy(self):
        tot, per = energy(self.U, self.K, self.P, self.label, self.g, len(self.Z))
        ekk = nuclear_energy(self.atoms[:, :3], self.Z, self.g.h64)
        return tot, per, ekk
