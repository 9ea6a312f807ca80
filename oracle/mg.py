"""Multigrid Poisson V-cycle oracle (TEST INFRASTRUCTURE ONLY).

New work: the reference has at most a two-grid correction
(``molecule.js:777-909``); the 4/5/7-level V-cycles of the target workloads are
specified here and pinned by mathematics, not by reference output.

All levels solve the h^2-scaled 7-point Dirichlet system
    A x = g,   (A x)_v = 6 x_v - sum_6 x_q,
with g = 2 pi h^2 rho on the finest level.  Level l has NN_l = NN / 2^l cells.

  smooth      x' = a x + (sum_6 x + g) / c      (a = 1/7, c = 7 is omega = 6/7)
  residual    r = g - (6 x - sum_6 x)
  restrict4   g_c = 4 * FW(r), FW = 27-point full weighting (8,4,2,1)/64;
              the factor 4 = (2h)^2 / h^2 keeps the scaled form on the coarse level
  prolong     trilinear interpolation, applied k-axis first, then j, then i
  exact_solve FP64 diagonalisation by the type-I discrete sine transform
  vcycle      nu1 pre-sweeps, residual, restrict, recurse (zero initial guess),
              prolong + correct (x += s * P e with s = +1), nu2 post-sweeps

The molecule.js two-grid correction (``molecule.js:777-909``) is the
special case (two levels, coarse rhs sign -1, 10 damped sweeps from zero in
place of the exact solve, correction scale s = +0.5); ``twogrid_s2`` states it.

Parity pinning (tests/test_oracle_mg.py):
  restrict4    weights sum to 64/64 and reproduce constants/affine fields (x4)
  prolong      exact on affine fields; P^T = 8 FW (adjoint identity)
  exact_solve  equals scipy.sparse.linalg.spsolve of A on random g
  vcycle       converges to spsolve's solution; contraction < 0.2 per cycle
  twogrid_s2   the correction has the opposite sign of the true error (Q2)
"""
from __future__ import annotations

import numpy as np
from scipy.fft import dstn, idstn

from .s1 import _interior, _nb, sum6, smooth


def residual(x, g, dtype):
    """r = g - (6 x - sum_6 x) on the interior (same shape as g)."""
    I = _interior(x.shape[0])
    return g - (dtype(6.0) * x[I] - sum6(x))


_FACES = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
_EDGES = [(1, 1, 0), (1, -1, 0), (-1, 1, 0), (-1, -1, 0),
          (1, 0, 1), (1, 0, -1), (-1, 0, 1), (-1, 0, -1),
          (0, 1, 1), (0, 1, -1), (0, -1, 1), (0, -1, -1)]
_CORNERS = [(1, 1, 1), (1, 1, -1), (1, -1, 1), (1, -1, -1),
            (-1, 1, 1), (-1, 1, -1), (-1, -1, 1), (-1, -1, -1)]


def restrict4(r_full, dtype):
    """Coarse right-hand side 4 * FW(r) at coarse interior nodes.

    ``r_full`` is a fine (S,S,S) array with zero faces.  Sum order: centre*8,
    faces*4 (order of _FACES), edges*2 (order of _EDGES), corners (order of
    _CORNERS); then * (4/64) = * 0.0625 (exact power of two).
    Returns the coarse (Sc,Sc,Sc) array with zero faces.
    """
    S = r_full.shape[0]
    NN = S - 1
    NC = NN // 2
    Sc = NC + 1
    idx = 2 * np.arange(1, NC)

    def at(di, dj, dk):
        return r_full[np.ix_(idx + di, idx + dj, idx + dk)]

    fsum = at(*_FACES[0])
    for o in _FACES[1:]:
        fsum = fsum + at(*o)
    esum = at(*_EDGES[0])
    for o in _EDGES[1:]:
        esum = esum + at(*o)
    csum = at(*_CORNERS[0])
    for o in _CORNERS[1:]:
        csum = csum + at(*o)
    val = dtype(8.0) * at(0, 0, 0) + dtype(4.0) * fsum + dtype(2.0) * esum + csum
    out = np.zeros((Sc, Sc, Sc), dtype)
    out[_interior(Sc)] = val * dtype(0.0625)
    return out


def _interp_axis(e, axis, dtype):
    """Linear interpolation of e (Sc nodes along axis) to 2*(Sc-1)+1 nodes."""
    Sc = e.shape[axis]
    S = 2 * (Sc - 1) + 1
    shape = list(e.shape)
    shape[axis] = S
    out = np.zeros(shape, dtype)
    even = [slice(None)] * 3
    even[axis] = slice(0, S, 2)
    out[tuple(even)] = e
    odd = [slice(None)] * 3
    odd[axis] = slice(1, S, 2)
    lo = [slice(None)] * 3
    lo[axis] = slice(0, Sc - 1)
    hi = [slice(None)] * 3
    hi[axis] = slice(1, Sc)
    out[tuple(odd)] = dtype(0.5) * (e[tuple(lo)] + e[tuple(hi)])
    return out


def prolong(ec, dtype):
    """Trilinear prolongation of a coarse (Sc,Sc,Sc) array: along k, then j, then i."""
    t = _interp_axis(ec, 2, dtype)
    t = _interp_axis(t, 1, dtype)
    return _interp_axis(t, 0, dtype)


def prolong_correct(x, ec, dtype, scale=1.0):
    """x += scale * P ec on the fine interior (faces untouched)."""
    I = _interior(x.shape[0])
    corr = prolong(ec, dtype)[I]
    out = x.copy()
    if scale == 1.0:
        out[I] = x[I] + corr
    else:
        out[I] = x[I] + dtype(scale) * corr
    return out


def dst_eigs(m):
    p = np.arange(1, m + 1)
    return 2.0 - 2.0 * np.cos(p * np.pi / (m + 1))


def exact_solve(g_full, dtype):
    """Exact FP64 solve of A e = g on the coarsest grid (Dirichlet zero faces).

    e = S^{-1} diag(1/(l_a + l_b + l_c)) S g with S the type-I DST along each
    axis (l_p = 2 - 2 cos(p pi / (m+1))).  Result cast to ``dtype``.
    """
    Sc = g_full.shape[0]
    m = Sc - 2
    I = _interior(Sc)
    gh = dstn(g_full[I].astype(np.float64), type=1)
    lam = dst_eigs(m)
    den = lam[:, None, None] + lam[None, :, None] + lam[None, None, :]
    e = idstn(gh / den, type=1)
    out = np.zeros((Sc, Sc, Sc), dtype)
    out[I] = e.astype(dtype)
    return out


def vcycle(x, g, levels, nu1, nu2, a, c, dtype):
    """One V-cycle on ``levels`` levels; x full (S,S,S), g interior rhs.

    levels == 1 means the exact solve on this grid.
    """
    S = x.shape[0]
    if levels == 1:
        gf = np.zeros((S, S, S), np.float64)
        gf[_interior(S)] = g
        return exact_solve(gf, dtype)
    for _ in range(nu1):
        x = smooth(x, g, a, c, dtype)
    r = np.zeros_like(x)
    r[_interior(S)] = residual(x, g, dtype)
    gc_full = restrict4(r, dtype)
    Sc = gc_full.shape[0]
    ec = vcycle(np.zeros((Sc, Sc, Sc), dtype), gc_full[_interior(Sc)], levels - 1, nu1, nu2, a, c, dtype)
    x = prolong_correct(x, ec, dtype)
    for _ in range(nu2):
        x = smooth(x, g, a, c, dtype)
    return x


def twogrid_s2(P, g, dtype, a=0.3333, c=9.0, coarse_sweeps=10, scale=0.5):
    """molecule.js two-grid correction restated in the scaled form.

    molecule.js forms r = -2 pi rho - lap P (molecule.js:777-803), restricts it
    (806-839), runs 10 damped sweeps E' = 0.3333 E + (sum E + 4h^2 f)/9 from
    E = 0 (842-867, 2893-2901) and adds 0.5 * trilinear(E) (870-909).  In the
    scaled form that is coarse rhs = -g_c, i.e. E approximates -e (sign
    inverted), then P += 0.5 * P E.
    """
    S = P.shape[0]
    r = np.zeros_like(P)
    r[_interior(S)] = residual(P, g, dtype)
    gc = restrict4(r, dtype)
    Sc = gc.shape[0]
    Ic = _interior(Sc)
    rhs = -gc[Ic]
    E = np.zeros((Sc, Sc, Sc), dtype)
    for _ in range(coarse_sweeps):
        E = smooth(E, rhs, a, c, dtype)
    return prolong_correct(P, E, dtype, scale)


def spsolve_poisson(g):
    """Reference direct solve of A x = g (interior m^3) with scipy, FP64."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    m = g.shape[0]
    T = sp.diags([-np.ones(m - 1), 2 * np.ones(m), -np.ones(m - 1)], [-1, 0, 1])
    I = sp.identity(m)
    A = (sp.kron(sp.kron(T, I), I) + sp.kron(sp.kron(I, T), I) + sp.kron(sp.kron(I, I), T)).tocsc()
    return spl.spsolve(A, g.astype(np.float64).ravel()).reshape(g.shape)
