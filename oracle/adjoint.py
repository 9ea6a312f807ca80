"""Energy of the fixed-step S1 trajectory as a differentiable function of the
atom positions (TEST INFRASTRUCTURE ONLY).

E(R) = T + V_eK + V_ee after: frozen-label trial init (K(R), U0 = trial of
each voxel's own label's atom), per-domain normalise, ``bootstrap`` Poisson
steps from P = 0, then ``n_steps`` steps of the cloud/solver.py loop body
(Poisson every ``poisson_interval`` steps, ITP, normalise every
``norm_interval``) with the labels held fixed (no boundary rule).  Written
in torch FP64 on the CPU so autograd gives an independent dE/dR for the
hand-written HIP adjoint.

Pins (tests/test_oracle_adjoint.py): trajectory_energy against the NumPy
oracle's passes on the same trajectory (oracle/s1.py init_trial, normalize,
jacobi, itp_update, energy and oracle/mg.py vcycle), Jacobi and multigrid
mode, to 1e-12 relative -- so _restrict4, _prolong, _vcycle and
_exact_solve are checked through it; its autograd gradient against central
differences of itself (delta = 1e-4 bohr) to 1e-7 Ha/bohr; the saved-start
map (``start``) by composition with the cold map (bit for bit) and by
central differences of itself.

Positions R are fractional grid units; gradients returned per bohr.
The Poisson step is either the cloud/solver.py Jacobi step (``mode='jacobi'``:
two sweeps of P' = (sum6 P + 2 pi h^2 U^2)/6) or one multigrid V-cycle
(``mode='multigrid'``, smoother (a, c), nu1 = nu2, exact coarsest solve).
"""
from __future__ import annotations

import math

import numpy as np
import torch

INV_SQRT_PI = 0.5641895835


def _sum6(P):
    return (P[2:, 1:-1, 1:-1] + P[:-2, 1:-1, 1:-1] + P[1:-1, 2:, 1:-1] + P[1:-1, :-2, 1:-1]
            + P[1:-1, 1:-1, 2:] + P[1:-1, 1:-1, :-2])


def _pad(inner):
    return torch.nn.functional.pad(inner, (1, 1, 1, 1, 1, 1))


def _smooth(P, rhs, a, c):
    new = (_sum6(P) + rhs) / c
    if a != 0:
        new = a * P[1:-1, 1:-1, 1:-1] + new
    return _pad(new)


def _restrict4(r):
    """4 * full weighting of a zero-face fine field onto coarse interior nodes."""
    S = r.shape[0]
    NC = (S - 1) // 2
    w1 = torch.tensor([1.0, 2.0, 1.0], dtype=r.dtype, device=r.device)
    w = (w1[:, None, None] * w1[None, :, None] * w1[None, None, :]) / 64.0
    # window J of the stride-2 convolution of r[1:, 1:, 1:] is centred on fine node 2(J+1),
    # i.e. coarse interior node I = J + 1
    out = torch.nn.functional.conv3d(r[None, None, 1:, 1:, 1:], w[None, None], stride=2)[0, 0]
    return 4.0 * out[: NC - 1, : NC - 1, : NC - 1]


def _prolong(ec):
    """Trilinear prolongation of a full coarse array (Sc^3) to (2Sc-1)^3."""
    def axis(e, d):
        Sc = e.shape[d]
        lo = e.narrow(d, 0, Sc - 1)
        hi = e.narrow(d, 1, Sc - 1)
        mid = 0.5 * (lo + hi)
        st = torch.stack([lo, mid], dim=d + 1)
        shape = list(e.shape)
        shape[d] = 2 * (Sc - 1)
        st = st.reshape(shape)
        return torch.cat([st, e.narrow(d, Sc - 1, 1)], dim=d)
    return axis(axis(axis(ec, 2), 1), 0)


def _exact_solve(gin):
    m = gin.shape[0]
    p = torch.arange(1, m + 1, dtype=gin.dtype, device=gin.device)
    Smat = math.sqrt(2.0 / (m + 1)) * torch.sin(math.pi * p[:, None] * p[None, :] / (m + 1))
    lam = 2.0 - 2.0 * torch.cos(math.pi * p / (m + 1))
    den = lam[:, None, None] + lam[None, :, None] + lam[None, None, :]
    t = torch.einsum("ap,pjk->ajk", Smat, gin)
    t = torch.einsum("bq,aqk->abk", Smat, t)
    t = torch.einsum("cr,abr->abc", Smat, t)
    t = t / den
    t = torch.einsum("ap,pjk->ajk", Smat, t)
    t = torch.einsum("bq,aqk->abk", Smat, t)
    t = torch.einsum("cr,abr->abc", Smat, t)
    return _pad(t)


def _vcycle(x, g, levels, nu, a, c):
    if levels == 1:
        return _exact_solve(g)
    for _ in range(nu):
        x = _smooth(x, g, a, c)
    r = _pad(g - (6.0 * x[1:-1, 1:-1, 1:-1] - _sum6(x)))
    gc = _restrict4(r)
    Sc = gc.shape[0] + 2
    ec = _vcycle(torch.zeros((Sc,) * 3, dtype=x.dtype, device=x.device), gc, levels - 1, nu, a, c)
    x = x + _pad(_prolong(ec)[1:-1, 1:-1, 1:-1])
    for _ in range(nu):
        x = _smooth(x, g, a, c)
    return x


def trajectory_energy(R, rows, label, NN, screen, n_steps, bootstrap=0, poisson_interval=2, norm_interval=20,
                      dv=None, mode="jacobi", mg_levels=3, mg_nu=2, mg_a=1.0 / 7.0, mg_c=7.0, start=None,
                      state_out=None, device=None):
    """E(R) for positions R (torch, (n,3), grid units) with Z, rc from ``rows``
    and frozen ``label`` (S^3 int array).  Returns (E, T, VeK, Vee) tensors.

    ``start`` = (U0, P0) (S^3 arrays): the map starts from these constants
    instead of the normalised trial and the bootstrap Poisson steps (the
    library's RQ_ADJ_START_SAVED); only K(R) then depends on R.
    ``state_out`` (a dict) receives the final U and P as numpy arrays.
    ``device``: where the tensors live (default: the CPU; the same FP64
    arithmetic on a GPU makes full-size grids affordable)."""
    dt = torch.float64
    dev = torch.device("cpu") if device is None else torch.device(device)
    S = NN + 1
    h = screen / NN
    h2, h3 = h * h, h ** 3
    Z = torch.tensor(np.asarray(rows)[:, 3], dtype=dt, device=dev)
    rc = torch.tensor(np.asarray(rows)[:, 4], dtype=dt, device=dev)
    n_elec = int((np.asarray(rows)[:, 3] > 0).sum())
    if dv is None:
        dv = 0.12 if n_elec <= 30 else 0.03
    hd, dtv = 0.5 * dv, dv * h2
    lab = torch.tensor(np.asarray(label).reshape(S, S, S)[1:-1, 1:-1, 1:-1].astype(np.int64), device=dev)
    x = torch.arange(1, S - 1, dtype=dt, device=dev) * h
    Kin = torch.zeros((S - 2,) * 3, dtype=dt, device=dev)
    U0in = torch.zeros((S - 2,) * 3, dtype=dt, device=dev)
    for n in range(len(Z)):
        if not Z[n] > 0:
            continue
        dx = (x - R[n, 0] * h)[:, None, None]
        dy = (x - R[n, 1] * h)[None, :, None]
        dz = (x - R[n, 2] * h)[None, None, :]
        r2 = dx * dx + dy * dy + dz * dz
        if rc[n] > 0:
            r = torch.clamp(torch.sqrt(r2 + 0.04 * h2), min=float(rc[n]))
        else:
            r = torch.sqrt(r2 + h2)
        Kin = Kin + Z[n] / r
        ut = Z[n] * Z[n] * INV_SQRT_PI * torch.exp(-Z[n] * r)
        U0in = torch.where(lab == n, ut, U0in)
    K = _pad(Kin)
    nd = len(Z)

    def normalize(U):
        ui = U[1:-1, 1:-1, 1:-1]
        norms = torch.zeros(nd, dtype=dt, device=dev).index_add(0, lab.reshape(-1), (ui * ui * h3).reshape(-1))
        ok = (Z > 0) & (norms > 1e-20)
        sc = torch.where(ok, torch.sqrt(torch.where(ok, Z / torch.where(ok, norms, torch.ones_like(norms)),
                                                     torch.ones_like(norms))), torch.ones_like(norms))
        return _pad(ui * sc[lab])

    def nbr_same(axis, sign):
        L = torch.tensor(np.asarray(label).reshape(S, S, S).astype(np.int64), device=dev)
        sl = [slice(1, -1)] * 3
        sl[axis] = slice(2, None) if sign > 0 else slice(0, -2)
        return L[tuple(sl)] == lab

    same = {(a_, s_): nbr_same(a_, s_) for a_ in range(3) for s_ in (1, -1)}

    def nb(U, axis, sign):
        sl = [slice(1, -1)] * 3
        sl[axis] = slice(2, None) if sign > 0 else slice(0, -2)
        return U[tuple(sl)]

    def itp(U, P):
        c = U[1:-1, 1:-1, 1:-1]
        lap = -6.0 * c
        for a_ in range(3):
            for s_ in (1, -1):
                lap = lap + torch.where(same[(a_, s_)], nb(U, a_, s_), c)
        new = c + hd * lap + dtv * (K[1:-1, 1:-1, 1:-1] - 2.0 * P[1:-1, 1:-1, 1:-1]) * c
        return _pad(new)

    def poisson(P, U):
        rhs = 2.0 * math.pi * U[1:-1, 1:-1, 1:-1] ** 2 * h2
        if mode == "jacobi":
            for _ in range(2):
                P = _smooth(P, rhs, 0.0, 6.0)
            return P
        return _vcycle(P, rhs, mg_levels, mg_nu, mg_a, mg_c)

    if start is None:
        U = normalize(_pad(U0in))
        P = torch.zeros((S,) * 3, dtype=dt, device=dev)
        for _ in range(bootstrap):
            P = poisson(P, U)
    else:
        U = torch.tensor(np.asarray(start[0], np.float64).reshape(S, S, S), device=dev)
        P = torch.tensor(np.asarray(start[1], np.float64).reshape(S, S, S), device=dev)
    for step in range(1, n_steps + 1):
        if step % poisson_interval == 0:
            P = poisson(P, U)
        U = itp(U, P)
        if step % norm_interval == 0:
            U = normalize(U)
    u = U[1:-1, 1:-1, 1:-1]
    T = 0.0
    for a_ in range(3):
        T = T + torch.where(same[(a_, 1)], nb(U, a_, 1) - u, torch.zeros_like(u)) ** 2
    T = 0.5 * h * T.sum()
    VeK = -(K[1:-1, 1:-1, 1:-1] * u * u).sum() * h3
    Vee = (P[1:-1, 1:-1, 1:-1] * u * u).sum() * h3
    if state_out is not None:
        state_out["U"] = U.detach().cpu().numpy().copy()
        state_out["P"] = P.detach().cpu().numpy().copy()
    return T + VeK + Vee, T, VeK, Vee


def energy_and_gradient(rows, label, NN, screen, n_steps, **kw):
    """(E, dE/dR per bohr as an (n,3) array) by autograd; R from ``rows``."""
    R = torch.tensor(np.asarray(rows)[:, :3], dtype=torch.float64, device=kw.get("device"), requires_grad=True)
    E, *_ = trajectory_energy(R, rows, label, NN, screen, n_steps, **kw)
    E.backward()
    h = screen / NN
    return float(E.detach()), R.grad.cpu().numpy() / h


def frozen_labels(rows, NN, screen):
    """Voronoi-by-trial labels at the given geometry (cloud/solver.py:36-53 argmax)."""
    from . import s1
    g = s1.Grid(NN, screen, 0.12, np.float64)
    _, lab, _, _ = s1.init_trial(rows, g)
    return lab
