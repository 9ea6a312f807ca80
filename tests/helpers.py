"""Shared helpers for the parity tests (host side only)."""
from __future__ import annotations

import numpy as np

from realmolecule_amd.builders import h2o_single, to_config_atoms


def small_h2o(NN=32, h=0.25):
    """Single H2O scaled onto a small grid (fractional positions)."""
    rows, els, screen = h2o_single(NN, h)
    return rows, els, screen


def three_atoms(NN=24):
    """Three well-separated domains (H, O-like with rc, H) on a small grid."""
    c = NN / 2
    rows = np.array([[c - 4.3, c + 0.7, c + 0.2, 1.0, 0.0],
                     [c + 0.4, c - 0.3, c - 0.6, 2.0, 0.4],
                     [c + 4.1, c + 1.9, c + 0.9, 1.0, 0.0]])
    return rows


def make_solver(rows, NN, screen, **cfg):
    from realmolecule_amd.solver import RealQMSolver
    config = {"gridN": NN, "screen": screen, "atoms": to_config_atoms(rows)}
    config.update(cfg)
    return RealQMSolver(config)


def cube(flat, S):
    return np.asarray(flat).reshape(S, S, S)


def ulp_diff(a, b):
    """Max difference in units of the last place between two float32 arrays."""
    a = np.ascontiguousarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.ascontiguousarray(b, np.float32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, -(a & 0x7fffffff), a)
    b = np.where(b < 0, -(b & 0x7fffffff), b)
    return int(np.abs(a - b).max()) if a.size else 0


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
