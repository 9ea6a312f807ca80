#!/usr/bin/env python3
"""Throughput of the RealQM grid-solver step on one MI355X.

Workload (default): 1000 H2O at 1 g/cm^3 (seed 4, O-O >= 2.5 A) in a
58.58-bohr cube, 3000 domains; 512^3 FP32 grid (h = 0.1144 bohr), 7-level
multigrid V(2,2) per ITP step, per-domain normalise every 20 steps,
nuclear-score boundary rule every 100 steps, FP64 (fixed-point) reductions.
A "step" is one ITP update plus one V-cycle.  Inputs are resident in HBM
before the timed region.

Prints ONE JSON line: value = Gvoxel-updates/s = (NN-1)^3 / ms_per_step,
where ms_per_step is the timed window's wall time per step with the
normalisation and the score rule counted once per 20 and per 100 steps
(their per-call device time, measured in the same run, added for the calls
the window missed or removed for the ones it over-counts); ``roofline`` for
the dominant fine kernel (algorithmic bytes per launch / average launch
time from HIP events on the solver's stream, plus measured HBM traffic from
rocprofv3 FETCH_SIZE / WRITE_SIZE); ``cpu_baseline`` (the C + OpenMP
restatement on all host cores, same fields); ``forces`` (adjoint against
central differences, on the 10-water 192^3 workload by default).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gvoxel-updates/s & %HBM roofline; batched adjoint forces/sec vs plain at ΔE≤1e-6"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per fine interior voxel and launch (each array counted
# once) of the fine-level passes.  Residual + restriction is one pass,
# prolongation + correction one, and the last post-sweep is fused with the
# ITP update.
BYTES_PER_VOXEL = {
    "itp": 18.0,         # U, K, P (f32) + label (u16) in; U' out  (steps without a V-cycle)
    "smooth": 12.0,      # P, U in; P' out
    "restrict": 8.5,     # P, U in; coarse rhs out (1/8 of 4 B)
    "prolong": 8.5,      # P in and out, coarse correction in (1/8)
    "smooth_itp": 22.0,  # last post-sweep + ITP: P, U, K, label in; P', U' out
    # pre-sweep pair + residual + restriction in one pass: two sweeps (2 x 12)
    # and the residual + restriction (8.5), each counted as its own unit
    "smooth_restrict": 32.5,
    # correction + post-sweep pair + ITP update in one pass: 8.5 + 2 x 12 + 18
    "post_itp": 50.5,
}

# Fine-level kernel of each family, as rocprofv3 names it, for the HBM
# traffic counters (FETCH_SIZE / WRITE_SIZE per dispatch).
# Several kernels run on every level; the fine level's dispatches are the
# ones within a factor 2 of the family's largest (the next level is 1/8).
PMC_KERNEL = {"smooth": ("k_smooth<true>", "k_smooth2_march<true"), "smooth_itp": ("k_smooth_itp",),
              "smooth_restrict": ("k_smooth2_resres_march",), "post_itp": ("k_post_march",),
              "itp": ("k_itp(",), "restrict": ("k_resres_march<true", "k_resres_tiled<true"),
              "prolong": ("k_prolong_correct", "k_prolong_smooth")}
COPY_MIB = 512  # plain device copy of known size that calibrates the counters


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def workload(name):
    from realmolecule_amd import builders as B
    if name == "c2":
        w = B.c2_waters()
        w["desc"] = "10 H2O (seed 1, O-O>=2.6A) in 12A cube, 192^3 FP32, 5-level V-cycle/step"
    elif name == "c1":
        w = B.c1_h2o()
        w["desc"] = "single H2O, 96^3 FP32 (h=0.1), 4-level V-cycle/step"
    elif name == "c4":
        w = B.c4_peptide()
        w["desc"] = "tri-glycine + 9 H2O (seed 3), 256^3 FP32, 5-level V-cycle/step"
    elif name == "c5":
        w = B.c5_waters()
        w["desc"] = "1000 H2O (seed 4, O-O>=2.5A) in 58.58 bohr cube, 512^3 FP32, 7-level V-cycle/step"
    elif name == "mf":
        w = {"mf": True, "cfg": B.mf_h2o(),
             "desc": "mol_fast.js H2O (O Z=3 rc=0.5 occupation 3, two H), 200^3 FP32, 3 electrons, "
                     "fused u/w/P update per step, normalise every 20, 2-grid per electron every 50"}
    else:
        raise SystemExit(f"unknown workload {name}")
    return w


def make_solver(w, graph=True):
    from realmolecule_amd.solver import RealQMSolver
    from realmolecule_amd.builders import to_config_atoms
    cfg = {"gridN": w["NN"], "screen": w["screen"], "atoms": to_config_atoms(w["rows"], w["els"]),
           "poisson": "multigrid", "mg_levels": w["mg_levels"], "mg_nu1": 2, "mg_nu2": 2,
           "bootstrap_poisson": 10, "graph": graph}
    return RealQMSolver(cfg)


def converge_steps(w, bootstrap, tol=1e-6, chunk=100, max_steps=60000):
    """Steps of the frozen-label solve (multigrid V-cycle every step) after
    which the electronic energy changes by <= tol relative over one more
    chunk of steps.  Returns (steps, relative change, energy)."""
    from realmolecule_amd.solver import RealQMSolver
    from realmolecule_amd.builders import to_config_atoms
    s = RealQMSolver({"gridN": w["NN"], "screen": w["screen"], "atoms": to_config_atoms(w["rows"], w["els"]),
                      "poisson": "multigrid", "mg_levels": w["mg_levels"], "bootstrap_poisson": bootstrap,
                      "boundary": "off", "graph": True})
    s.initialize()
    prev, d, E, T = None, 1.0, 0.0, 0
    while T < max_steps:
        s.step(T, chunk, poisson_interval=1)
        T += chunk
        s._compute_energy()
        E = s.E_T + s.E_eK + s.E_ee
        if prev is not None:
            d = abs(E - prev) / abs(E)
            if d <= tol:
                break
        prev = E
    s.close()
    return T, d, E


def forces_bench(w, bootstrap, n_batch, fd_delta, fd_components, log_fn, conv_tol=1e-6):
    """Forces/sec at converged energy.  The solve is the frozen-label
    trajectory (trial at the labels, normalise, bootstrap V-cycles, then the
    cloud/solver.py loop body with a V-cycle per step) run until its energy
    changes by <= conv_tol relative over 100 more steps.  Plain path: central
    differences (delta = fd_delta bohr) of that full solve, 6N+1 solves per
    force vector, timed on a sample of components; adjoint: one forward solve
    plus the reverse sweep with checkpointed recomputation; batched: n_batch
    geometries in one batch (rq_batch_adjoint_gradient).  Both paths evaluate
    the same E(R)."""
    import ctypes as C
    from realmolecule_amd.adjoint import FrozenSolver
    NN, screen, rows = w["NN"], w["screen"], w["rows"]
    T, dconv, Econv = converge_steps(w, bootstrap, conv_tol)
    log_fn(f"[forces] converged: {T} steps, dE/E over the last 100 steps {dconv:.2e}")
    kw = dict(bootstrap=bootstrap, checkpoint=0, poisson_interval=1, norm_interval=20, poisson="multigrid",
              mg_levels=w["mg_levels"])
    fz = FrozenSolver(rows, NN, screen, T, **kw)
    pos = rows[:, :3] * fz.h
    n = len(rows)
    t0 = time.perf_counter()
    E_adj, g = fz.energy_and_gradient(pos)
    t_adj = time.perf_counter() - t0
    log_fn(f"[forces] adjoint: {t_adj:.3f} s for {3 * n} components ({T} steps)")
    # plain path on a sample of components: each is two full solves
    t0 = time.perf_counter()
    E0 = fz.energy(pos)
    t_e0 = time.perf_counter() - t0
    # the reference's own force at the converged state: Hellmann-Feynman
    # (cloud/solver.py:282-315) on the final U of that solve
    hf = (C.c_double * (3 * n))()
    fz.s._chk(fz.lib.rq_forces_hf(fz.hdl, 0, n, hf))
    F_hf = np.frombuffer(hf, np.float64).reshape(n, 3).copy()
    from realmolecule_amd.solver import nuclear_forces
    F_el_adj = -(g + nuclear_forces(rows[:, :3].astype(np.float32), rows[:, 3], fz.h))
    bare = rows[:, 4] <= 0
    comps = []
    for a in list(np.nonzero(bare)[0][:1]) + list(np.nonzero(~bare)[0][:1]):
        for c in range(3):
            comps.append((int(a), c))
    comps = comps[:fd_components] if fd_components <= len(comps) else \
        [(a, c) for a in range(n) for c in range(3)][:fd_components]
    F_fd = {}
    t0 = time.perf_counter()
    for a, c in comps:
        F_fd[(a, c)] = _fd_component(fz, pos, a, c, fd_delta)
    t_fd_sample = time.perf_counter() - t0
    t_solve = (t_fd_sample + t_e0) / (2 * len(comps) + 1)
    t_fd = (6 * n + 1) * t_solve
    diffs = [abs(-g[a, c] - F_fd[(a, c)]) for a, c in comps]
    log_fn(f"[forces] central differences: {2 * len(comps) + 1} solves in {t_fd_sample + t_e0:.2f} s, "
           f"{t_solve:.3f} s per solve, {t_fd:.1f} s per force vector (6N+1 = {6 * n + 1} solves)")
    fz.close()
    # batched: n_batch geometries (seeded 0.05 bohr jitter, each with its own
    # trial labels) as one batch: every geometry's frozen map and reverse
    # sweep in one launch per pass (rq_batch_adjoint_gradient), checkpoints
    # every 64 steps, allocations included
    from realmolecule_amd.batch import Batch
    from realmolecule_amd.adjoint import make_spec
    rng = np.random.default_rng(7)
    h = screen / NN
    geo_rows = []
    for _ in range(n_batch):
        r = rows.copy()
        r[:, :3] = (pos + rng.normal(0, 0.05, pos.shape)) / h
        geo_rows.append(r)
    b = Batch([(r, NN, screen) for r in geo_rows], mg_levels=w["mg_levels"], graph=False)
    b.initialize(bootstrap=0)
    t0 = time.perf_counter()
    E3b, grads_b = b.adjoint_gradient(make_spec(T, bootstrap, 64, 1, 20))
    t_batch = time.perf_counter() - t0
    b.close()
    log_fn(f"[forces] batched adjoint: {n_batch} force vectors in {t_batch:.3f} s")
    hf_diff = np.abs(F_el_adj - F_hf)
    return {
        "workload": f"{w['desc']}; frozen-label solve of {bootstrap} bootstrap V-cycles + {T} steps "
                    f"(converged: dE/E {dconv:.2e} over the last 100 steps)",
        "n_atoms": n,
        "steps_to_converge": T,
        "converged_dE_rel_per_100_steps": dconv,
        "energy_Ha": E0,
        "energy_rel_diff_adjoint_vs_plain": abs(E_adj - E0) / abs(E0),
        "plain_fd_forces_per_s": 1.0 / t_fd,
        "plain_fd_solves": 6 * n + 1,
        "plain_fd_seconds_per_solve": t_solve,
        "plain_fd_timed_solves": 2 * len(comps) + 1,
        "adjoint_forces_per_s": 1.0 / t_adj,
        "batched_adjoint_forces_per_s": n_batch / t_batch,
        "batch": n_batch,
        "speedup_adjoint": t_fd / t_adj,
        "speedup_batched_adjoint": (n_batch / t_batch) * t_fd,
        "fd_delta_bohr": fd_delta,
        "fd_checked": [{"atom": a, "axis": c, "adjoint": -g[a, c], "fd": F_fd[(a, c)],
                        "rc": float(rows[a, 4])} for a, c in comps],
        "max_abs_force_diff_vs_fd_Ha_per_bohr": float(max(diffs)),
        "max_abs_force_Ha_per_bohr": float(np.abs(g).max()),
        "hf_electronic_vs_adjoint_rc0_atoms_max_abs": float(hf_diff[bare].max()) if bare.any() else None,
        "hf_electronic_vs_adjoint_rc0_atoms_rms": float(np.sqrt((hf_diff[bare] ** 2).mean())) if bare.any() else None,
    }


def bench_full_adjoint(wname, window, converge_max, fd_components, log_fn, tol=1e-6, chunk=100):
    """One full adjoint force vector on a large workload (all atoms), from a
    converged start.  The frozen-label solve (trial, normalise, 10 bootstrap
    V-cycles, then a V-cycle + ITP per step) is run in chunks of ``chunk``
    steps until the electronic energy changes by <= tol relative over one
    chunk (or ``converge_max`` steps); that state is stored
    (rq_adjoint_save_start) and the differentiated map is the ``window``
    steps that follow it, at any positions (RQ_ADJ_START_SAVED).  Plain path:
    central differences (delta = 1e-3 bohr) of that map, 6N+1 solves per
    force vector, timed on ``fd_components`` components; the reference's own
    force (Hellmann-Feynman, cloud/solver.py:282-315) at the stored state for
    the rc = 0 atoms."""
    import ctypes as C
    from realmolecule_amd import _lib as L
    from realmolecule_amd.adjoint import FrozenSolver
    from realmolecule_amd.solver import nuclear_forces
    w = workload(wname)
    NN, screen, rows = w["NN"], w["screen"], w["rows"]
    n = len(rows)
    fz = FrozenSolver(rows, NN, screen, chunk, bootstrap=10, checkpoint=16, poisson_interval=1, norm_interval=20,
                      poisson="multigrid", mg_levels=w["mg_levels"])
    pos = rows[:, :3] * fz.h
    ekk = fz._ekk(fz.rows)
    t0 = time.perf_counter()
    E_prev = fz.warm_start(pos, chunk) - ekk
    T, d = chunk, 1.0
    while T < converge_max:
        E = fz.energy(pos) - ekk
        fz.s._chk(fz.lib.rq_adjoint_save_start(fz.hdl))
        T += chunk
        d = abs(E - E_prev) / abs(E)
        E_prev = E
        if d <= tol:
            break
    t_conv = time.perf_counter() - t0
    log_fn(f"[adjoint] {wname}: start after {T} steps ({t_conv:.1f} s), dE/E over the last {chunk} steps {d:.2e}")
    hf = (C.c_double * (3 * n))()
    fz.s._chk(fz.lib.rq_forces_hf(fz.hdl, 0, n, hf))
    F_hf = np.frombuffer(hf, np.float64).reshape(n, 3).copy()
    fz.spec.n_steps = window
    t0 = time.perf_counter()
    E0 = fz.energy(pos)
    t_e0 = time.perf_counter() - t0
    t0 = time.perf_counter()
    E_adj, g = fz.energy_and_gradient(pos)
    t_adj = time.perf_counter() - t0
    log_fn(f"[adjoint] {wname}: {3 * n} components in {t_adj:.3f} s ({window}-step window; one solve {t_e0:.3f} s)")
    F_el_adj = -(g + nuclear_forces(rows[:, :3].astype(np.float32), rows[:, 3], fz.h))
    bare = rows[:, 4] <= 0
    comps = [(int(a), c) for a in list(np.nonzero(bare)[0][:1]) + list(np.nonzero(~bare)[0][:1]) for c in range(3)]
    # FP32 amplitudes put ~|E| 2^-24 of rounding noise on each solve's
    # energy; beyond ~1e-2 Ha/bohr over 2 delta central differences resolve
    # nothing (a 1000-water energy of -2e5 Ha: ~6 Ha/bohr)
    fd_noise = abs(E0) * 2.0 ** -24 / 2e-3
    if fd_noise > 1e-2:
        comps = []
    comps = comps[:fd_components]
    F_fd = {}
    t0 = time.perf_counter()
    for a, c in comps:
        F_fd[(a, c)] = _fd_component(fz, pos, a, c, 1e-3)
    t_fd_sample = time.perf_counter() - t0
    t_solve = (t_fd_sample + t_e0) / (2 * len(comps) + 1)
    t_fd = (6 * n + 1) * t_solve
    log_fn(f"[adjoint] {wname}: central differences {t_solve:.3f} s per solve, {t_fd:.0f} s per force vector"
           f"{'' if comps else f' (not evaluated: FP32 noise ~{fd_noise:.1f} Ha/bohr)'}")
    fz.close()
    hf_diff = np.abs(F_el_adj - F_hf)[bare]
    return {
        "metric": "full adjoint force vector, one GPU",
        "workload": f"{w['desc']}; {n} atoms; start converged by {T} frozen-label steps (dE/E {d:.2e} over "
                    f"the last {chunk}), differentiated map = the {window} steps after it",
        "n_atoms": n,
        "components": 3 * n,
        "start_steps": T,
        "start_dE_rel_per_chunk": d,
        "start_seconds": t_conv,
        "window_steps": window,
        "energy_Ha": E0,
        "energy_rel_diff_adjoint_vs_plain": abs(E_adj - E0) / abs(E0),
        "adjoint_seconds": t_adj,
        "adjoint_forces_per_s": 1.0 / t_adj,
        "adjoint_over_one_solve": t_adj / t_e0,
        "plain_fd_seconds_per_solve": t_solve,
        "plain_fd_solves": 6 * n + 1,
        "plain_fd_forces_per_s": 1.0 / t_fd,
        "speedup_adjoint": t_fd / t_adj,
        "fd_delta_bohr": 1e-3,
        "fd_noise_estimate_Ha_per_bohr": fd_noise,
        "fd_checked": [{"atom": a, "axis": c, "adjoint": -g[a, c], "fd": F_fd[(a, c)], "rc": float(rows[a, 4])}
                       for a, c in comps],
        "hf_vs_adjoint_rc0_atoms_max_abs": float(hf_diff.max()) if bare.any() else None,
        "hf_vs_adjoint_rc0_atoms_rms": float(np.sqrt((hf_diff ** 2).mean())) if bare.any() else None,
        "max_abs_force_Ha_per_bohr": float(np.abs(g).max()),
    }


def bench_relax(wname, iters, fd_iters, n_steps, log_fn, converge=2000):
    """Differentiable relaxation: L-BFGS-B (scipy) on adjoint gradients for
    ``iters`` iterations, against the same driver on central-difference
    gradients (6N+1 solves per gradient, delta = 1e-3 bohr) for ``fd_iters``
    iterations.  The energy is one fixed function of the positions: the
    ``n_steps`` frozen-label steps that follow a state converged by
    ``converge`` steps at the start geometry (checkpoints every 64 steps in
    the reverse sweep).  Reports wall time per iteration of each and the
    energies both reach after ``fd_iters`` iterations."""
    from realmolecule_amd.relax import relax
    w = workload(wname)
    NN, screen, rows = w["NN"], w["screen"], w["rows"]
    kw = dict(bootstrap=10, checkpoint=64, poisson_interval=1, poisson="multigrid", mg_levels=w["mg_levels"],
              converge=converge, ftol=0.0, gtol=1e-12)
    ad_short = relax(rows, NN, screen, n_steps=n_steps, max_iter=fd_iters, **kw)
    ad = relax(rows, NN, screen, n_steps=n_steps, max_iter=iters, **kw)
    log_fn(f"[relax] {wname}: adjoint {len(ad['grad_max'])} iterations in {ad['wall_s']:.1f} s, "
           f"E {ad['energies'][0]:.6f} -> {min(ad['energies']):.6f}")
    fd = relax(rows, NN, screen, n_steps=n_steps, max_iter=fd_iters, fd=True, fd_delta=1e-3, log=log_fn, **kw)
    log_fn(f"[relax] {wname}: central differences {len(fd['grad_max'])} iterations in {fd['wall_s']:.1f} s")
    it_ad = max(1, len(ad["grad_max"]))
    it_fd = max(1, len(fd["grad_max"]))
    return {
        "metric": "relaxation iterations per second, adjoint vs central-difference gradients",
        "workload": f"{w['desc']}; {len(rows)} atoms; energy = {n_steps} frozen-label steps after a start "
                    f"converged by {converge} steps at the start geometry; L-BFGS-B",
        "start_seconds": ad["start_s"],
        "adjoint": {"iterations": len(ad["grad_max"]), "evaluations": ad["evaluations"], "wall_s": ad["wall_s"],
                    "s_per_iteration": ad["wall_s"] / it_ad, "E_first": ad["energies"][0],
                    "E_min": min(ad["energies"]), "grad_max_first": ad["grad_max"][0] if ad["grad_max"] else None,
                    "grad_max_last": ad["grad_max"][-1] if ad["grad_max"] else None, "message": ad["message"]},
        "central_differences": {"iterations": len(fd["grad_max"]), "evaluations": fd["evaluations"],
                                "wall_s": fd["wall_s"], "s_per_iteration": fd["wall_s"] / it_fd,
                                "E_first": fd["energies"][0], "E_min": min(fd["energies"]), "fd_delta_bohr": 1e-3},
        "adjoint_same_iterations": {"iterations": len(ad_short["grad_max"]), "E_min": min(ad_short["energies"]),
                                    "wall_s": ad_short["wall_s"]},
        "speedup_per_iteration": (fd["wall_s"] / it_fd) / (ad["wall_s"] / it_ad),
    }


def _fd_component(fz, pos, a, c, delta):
    """Central difference of the full solve for one force component (the
    step divided by the difference of the stored FP32 grid positions)."""
    p = np.asarray(pos, np.float64).copy()
    p[a, c] = pos[a, c] + delta
    xp = fz.stored_position(p[a, c])
    ep = fz.energy(p)
    p[a, c] = pos[a, c] - delta
    xm = fz.stored_position(p[a, c])
    em = fz.energy(p)
    return -(ep - em) / (xp - xm)


def bench_hairpin():
    """The reference's recorded run, re-run: the bent solvated hairpin of
    cloud/run_hairpin.py (267 atoms, 300^3 cells, Jacobi Poisson, dynamics
    with forceScale 500, energy reports every 1000 steps) through
    RealQMSolver.run(10000), timed like the reference's ms_per_step (wall time
    of run() over its steps, host work included), next to the A100's 5.277
    ms/step for the same run (cloud/hairpin_results.json)."""
    from realmolecule_amd.builders import hairpin_solvated
    from realmolecule_amd.solver import RealQMSolver
    cfg = hairpin_solvated()
    cfg["graph"] = True
    s = RealQMSolver(cfg)
    s.set_fold_atoms(cfg["fold_atoms"])
    s.initialize()
    import contextlib
    t0 = time.perf_counter()
    with contextlib.redirect_stdout(sys.stderr):  # run() prints a line per report; stdout carries the JSON line
        snaps = s.run(cfg["steps"], report_interval=cfg["report_interval"])
    dt = time.perf_counter() - t0
    s.close()
    NN = cfg["gridN"]
    ms = dt / cfg["steps"] * 1e3
    a100 = 5.277
    log(f"[hairpin] {ms:.4f} ms/step against the A100's {a100} ms/step")
    return {
        "metric": "Gvoxel-updates/s (reference hairpin run, 300^3, dynamics on)",
        "value": round((NN - 1) ** 3 / (ms * 1e-3) / 1e9, 3), "unit": "Gvoxel-updates/s", "n_gpus": 1,
        "steps": cfg["steps"], "warmup": 0, "ms_per_step": round(ms, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": round(a100 / ms, 3), "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "bent solvated hairpin of cloud/run_hairpin.py: 87 protein atoms + 60 waters, "
                               "300^3 FP32 (h = 0.267 bohr), Jacobi Poisson (2 sweeps every 2 steps), dynamics every "
                               "500 steps with forceScale 500, reports every 1000", "parallelism": "single-gpu"},
        "baseline": {"ms_per_step": a100, "hardware": "NVIDIA A100 (cloud/hairpin_results.json)",
                     "note": "vs_baseline = A100 ms/step / MI355X ms/step on the same run"},
        "energy_Ha": snaps[-1]["E"], "fold_angle": snaps[-1]["fold_angle"],
    }


def bench_c3(steps, warmup, graph=True, sequential=True):
    """Batched geometry scan: 64 seeded conformers (H2O, formamide, ethanol,
    glycine) on grids of 64..160 cells, 4-level V(2,2) + ITP per step.  The
    batch engine advances all of them with one launch per pass over all of
    their tiles;
    the comparison runs the same molecules one after another, each on its own
    single-grid handle (the per-molecule path a scan would take)."""
    from realmolecule_amd import builders as B
    from realmolecule_amd.batch import Batch
    from realmolecule_amd.solver import RealQMSolver
    mols = B.c3_conformers()
    vox = sum((nn - 1) ** 3 for _, nn, _, _ in mols)
    b = Batch([(r, nn, sc) for r, nn, sc, _ in mols], mg_levels=4, graph=graph)
    b.initialize(bootstrap=10)
    b.step(0, warmup)
    b.sync()
    t0 = time.perf_counter()
    b.step(warmup, steps)
    b.sync()
    dt_b = time.perf_counter() - t0
    b.close()
    log(f"[c3] batch: {dt_b / steps * 1e3:.3f} ms/step for {len(mols)} molecules, {vox * steps / dt_b / 1e9:.2f} Gvox/s")
    seq_steps = max(20, steps // 5)
    dt_s = 0.0
    for r, nn, sc, name in (mols if sequential else []):
        from realmolecule_amd.builders import to_config_atoms
        s = RealQMSolver({"gridN": nn, "screen": sc, "atoms": to_config_atoms(r), "poisson": "multigrid",
                          "mg_levels": 4, "bootstrap_poisson": 10, "graph": True})
        s.initialize()
        s.step(0, 20, poisson_interval=1)
        s.sync()
        t0 = time.perf_counter()
        s.step(20, seq_steps, poisson_interval=1)
        s.sync()
        dt_s += time.perf_counter() - t0
        s.close()
    log(f"[c3] one handle per molecule, in turn: {dt_s / seq_steps * 1e3:.3f} ms per step of all molecules")
    return {
        "metric": "Gvoxel-updates/s (64-conformer batch)",
        "value": round(vox * steps / dt_b / 1e9, 3),
        "unit": "Gvoxel-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(dt_b / steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "64 conformers (seed 2) of H2O/formamide/ethanol/glycine, grids 64..160 cells, "
                               "4-level V(2,2) + ITP per step", "molecules": len(mols), "interior_voxels": vox,
                   "parallelism": "single-gpu"},
        "batch_graph_replay": graph,
        "sequential_ms_per_step": round(dt_s / seq_steps * 1e3, 5) if sequential else None,
        "batch_speedup": round((dt_s / seq_steps) / (dt_b / steps), 3) if sequential else None,
    }


def bench_c3_forces(n_steps, checkpoint, log_fn):
    """Batched adjoint forces on the 64-conformer scan: the frozen-label map
    (trial, normalise, 10 bootstrap V-cycles, n_steps steps of a 4-level
    V(2,2) + ITP) of every conformer and its reverse sweep in one launch per
    pass (rq_batch_adjoint_gradient), against the same map and adjoint on one
    handle per conformer in turn, and against central differences (6N+1
    solves per conformer, extrapolated from each conformer's timed solve)."""
    from realmolecule_amd import builders as B
    from realmolecule_amd.adjoint import FrozenSolver, make_spec
    from realmolecule_amd.batch import Batch
    mols = B.c3_conformers()
    spec = make_spec(n_steps, bootstrap=10, checkpoint=checkpoint, poisson_interval=1, norm_interval=20)
    b = Batch([(r, nn, sc) for r, nn, sc, _ in mols], mg_levels=4, graph=False)
    b.initialize(bootstrap=0)
    b.adjoint_gradient(spec)  # warm: allocations
    b.sync()
    t0 = time.perf_counter()
    E3, grads = b.adjoint_gradient(spec)
    t_batch = time.perf_counter() - t0
    b.close()
    log_fn(f"[c3 forces] batched: {len(mols)} force vectors in {t_batch:.3f} s")
    t_seq = t_fd = 0.0
    max_diff = 0.0
    n_atoms = 0
    for m, (r, nn, sc, name) in enumerate(mols):
        fz = FrozenSolver(r, nn, sc, n_steps, bootstrap=10, checkpoint=checkpoint, poisson_interval=1,
                          poisson="multigrid", mg_levels=4)
        pos = r[:, :3] * fz.h
        fz.energy_and_gradient(pos)  # warm
        fz.s.sync()
        t0 = time.perf_counter()
        E, g = fz.energy_and_gradient(pos)
        t_seq += time.perf_counter() - t0
        t0 = time.perf_counter()
        fz.energy(pos)
        t_solve = time.perf_counter() - t0
        t_fd += (6 * len(r) + 1) * t_solve
        from realmolecule_amd.solver import nuclear_forces
        g_el = g + nuclear_forces(r[:, :3].astype(np.float32), r[:, 3], fz.h)
        max_diff = max(max_diff, float(np.abs(g_el - grads[m]).max()))
        n_atoms += len(r)
        fz.close()
    log_fn(f"[c3 forces] one handle per conformer: {t_seq:.3f} s; central differences {t_fd:.1f} s")
    n = len(mols)
    return {
        "metric": "batched adjoint force vectors/s (64-conformer scan)",
        "workload": f"64 conformers (seed 2) of H2O/formamide/ethanol/glycine, grids 64..160 cells, {n_atoms} atoms; "
                    f"frozen-label map of 10 bootstrap V-cycles + {n_steps} steps (4-level V(2,2) + ITP), "
                    f"checkpoints every {checkpoint} steps",
        "batched_adjoint_forces_per_s": n / t_batch,
        "sequential_adjoint_forces_per_s": n / t_seq,
        "plain_fd_forces_per_s": n / t_fd,
        "batch_speedup_over_sequential_adjoint": t_seq / t_batch,
        "speedup_batched_adjoint_over_fd": t_fd / t_batch,
        "max_abs_dEdR_diff_batched_vs_single": max_diff,
        "seconds": {"batched": t_batch, "sequential_adjoint": t_seq, "fd_extrapolated": t_fd},
    }


def bench_mf(args):
    """mol_fast.js water: frames of 500 fused steps (the page's doSteps),
    inputs resident on the device; roofline of the fused update from HIP
    events around each launch."""
    from realmolecule_amd.molfast import MolFast
    w = workload("mf")
    host = MolFast(w["cfg"])
    NN, NE = host.NN, host.NE
    n3 = (NN - 1) ** 3
    log(f"[bench] mf: {w['desc']}; warmup {args.warmup} steps")
    host.do_steps(args.warmup)
    log(f"[bench] mf: timed {args.steps} steps in frames of 500")
    left = args.steps
    t0 = time.perf_counter()
    while left > 0:
        n = min(500, left)
        host.do_steps(n)  # returns after the frame's reduction is copied back
        left -= n
    dt = time.perf_counter() - t0
    ms_per_step = dt / args.steps * 1e3
    value = n3 * NE * args.steps / dt / 1e9
    log(f"[bench] mf: {ms_per_step:.4f} ms/step, {value:.2f} Gvoxel-electron updates/s, E={host.E:.6f}")
    host.set_timing(True)
    host.do_steps(args.timing_steps)
    ms, cnt = host.update_time()
    host.set_timing(False)
    avg_us = ms / cnt * 1e3
    bpn = 24.0 * NE + 4.0  # u, w, P in and out per electron, K in
    achieved = bpn * n3 / (avg_us * 1e-6) / 1e9
    roof = {"kernel": "mf_update", "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "traffic_unit": "bytes per launch", "alg_bytes_per_node": bpn, "alg_bytes_per_launch": bpn * n3,
            "avg_us": round(avg_us, 3), "launches": cnt}
    forces = host.compute_forces()
    E = host.E
    K = host.field("K")
    fields = [np.stack([host.field(f, m) for m in range(NE)]) for f in ("U", "W", "P")]
    host.close()
    if not args.no_pmc:
        log("[bench] mf: HBM traffic of the fused update: rocprofv3 FETCH_SIZE / WRITE_SIZE passes")
        traffic, info = pmc_traffic("mf", {"mf_update": ("k_mf_update",)})
        if traffic and "mf_update" in traffic:
            roof["traffic"] = round(traffic["mf_update"])
            roof["traffic_over_alg"] = round(traffic["mf_update"] / (bpn * n3), 4)
        roof["traffic_pmc"] = info
    cpu = None
    if not args.no_cpu:
        from oracle import molfast as O
        nuc = [O.Nucleus(n["i"], n["j"], n["k"], n["Z"], rc=n["rc"]) for n in w["cfg"]["USER_NUCLEI"]]
        mf = O.MolFast(nuc, NN=NN, screen=w["cfg"]["USER_SCREEN"], targets=w["cfg"]["USER_NORM_TARGETS"])
        U, W, P = fields
        steps = 2
        log(f"[bench] mf: cpu baseline, {steps} fused updates of the NumPy FP32 oracle")
        t0 = time.perf_counter()
        for _ in range(steps):
            U, W, P = O.update(mf, K, U, W, P)
        cdt = time.perf_counter() - t0
        cpu = {"value": round(n3 * NE * steps / cdt / 1e9, 6), "unit": "Gvoxel-electron updates/s", "cores": 1,
               "kind": "port", "sample": f"{steps} fused updates of the same {NN}^3, 3-electron fields (read back "
                                         f"from the device), NumPy FP32 restatement, {cdt:.1f} s"}
    return {
        "metric": "mol_fast fused per-electron steps: Gvoxel-electron updates/s & %HBM roofline",
        "value": round(value, 3), "unit": "Gvoxel-electron updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": w["desc"], "NN": NN, "electrons": NE, "interior_voxels": n3,
                   "steps_per_frame": 500, "parallelism": "single-gpu"},
        "roofline": roof, "cpu_baseline": cpu, "energy_Ha": E,
        "forces": [[round(x, 6) for x in f] for f in forces],
    }


def bench_pother(args, kind):
    """Per-domain potentials of molecule.js.  pd: one water, direct Pother
    (per domain: rho_other and one V-cycle on its persistent P_m each step,
    then the label gather).  sic: the 10-water grid on molecule.js's frame
    schedule (2 damped sweeps per step, 2-grid correction every 50, ITP on
    the frame potential), frames of 50 steps, each ending with the frame
    potential and, every 5th frame, the self-interaction correction of all
    30 domains."""
    import ctypes as C
    from realmolecule_amd import _lib as L
    from realmolecule_amd import builders as B
    from realmolecule_amd.solver import RealQMSolver
    if kind == "pd":
        NN = 192
        rows, els, screen = B.h2o_single(NN, 0.06)
        cfg = {"gridN": NN, "screen": screen, "atoms": B.to_config_atoms(rows, els), "poisson": "multigrid",
               "mg_levels": 5, "mg_nu1": 2, "mg_nu2": 2, "bootstrap_poisson": 4, "pother": "direct", "graph": True}
        desc = ("1 H2O, 192^3 (h = 0.06 bohr), direct per-domain potential: per step and domain rho_other + one "
                "5-level V(2,2) on the domain's persistent P_m, label gather, ITP")
        frame = None
    else:
        w = B.c2_waters()
        NN = w["NN"]
        cfg = {"gridN": NN, "screen": w["screen"], "atoms": B.to_config_atoms(w["rows"], w["els"]),
               "poisson": "twogrid", "jacobi_sweeps": 2, "smooth_a": 0.3333, "smooth_c": 9.0, "twogrid_interval": 50,
               "pother": "frame", "sic_interval": 5, "init": "s2", "energy": "s2", "boundary": "off",
               "bootstrap_poisson": 10, "graph": True}
        desc = ("10 H2O (seed 1), 192^3, molecule.js frame schedule: 2 damped sweeps/step + 2-grid every 50, ITP on "
                "the frame potential; frames of 50 steps, SIC of all 30 domains every 5th frame")
        frame = 50
    n3 = (NN - 1) ** 3
    s = RealQMSolver(cfg)
    s.initialize()
    fno = 0

    def run(first, n):
        nonlocal fno
        if frame is None:
            s.step(first, n, poisson_interval=1)
            return
        done = 0
        while done < n:
            m = min(frame, n - done)
            s.step(first + done, m, poisson_interval=1, norm_interval=20, boundary_interval=10 ** 9)
            done += m
            fno += 1
            s.pother_frame(fno)

    steps = args.steps if frame is None else max(frame, (args.steps // (5 * frame)) * 5 * frame)
    log(f"[bench] {kind}: {desc}; warmup {args.warmup}, timed {steps}")
    run(0, args.warmup if frame is None else frame * 5)
    s.sync()
    first = args.warmup if frame is None else frame * 5
    t0 = time.perf_counter()
    run(first, steps)
    s.sync()
    dt = time.perf_counter() - t0
    ms_per_step = dt / steps * 1e3
    value = n3 * steps / dt / 1e9
    log(f"[bench] {kind}: {ms_per_step:.4f} ms/step, {value:.2f} Gvoxel-updates/s")
    lib = s._lib
    s._chk(lib.rq_timing_enable(s._h, 1))
    run(first + steps, args.timing_steps if frame is None else frame * 5)
    nf = len(L.FAMILIES)
    ms = (C.c_double * nf)()
    cnt = (C.c_int64 * nf)()
    s._chk(lib.rq_timing_read(s._h, ms, cnt, nf))
    kernels = {}
    for f, name in enumerate(L.FAMILIES):
        if cnt[f]:
            avg = ms[f] / cnt[f] * 1e3
            ent = {"launches": int(cnt[f]), "avg_us": round(avg, 3), "total_ms": round(ms[f], 4)}
            if name in BYTES_PER_VOXEL:
                ent["achieved_GBs"] = round(BYTES_PER_VOXEL[name] * n3 / (avg * 1e-6) / 1e9, 1)
            kernels[name] = ent
    fine = {k: v for k, v in kernels.items() if "achieved_GBs" in v}
    dom = max(fine, key=lambda k: fine[k]["total_ms"])
    roof = {"kernel": dom, "bound": "hbm", "achieved": fine[dom]["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(fine[dom]["achieved_GBs"] / HBM_PEAK_GBS, 4), "traffic": None,
            "alg_bytes_per_voxel": BYTES_PER_VOXEL[dom]}
    s._compute_energy()
    E = s.E
    s.close()
    return {"metric": f"per-domain potential ({kind}): Gvoxel-updates/s & %HBM roofline", "value": round(value, 3),
            "unit": "Gvoxel-updates/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
            "ms_per_step": round(ms_per_step, 5), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "fp32", "data": "synthetic",
            "config": {"workload": desc, "NN": NN, "interior_voxels": n3, "parallelism": "single-gpu"},
            "roofline": roof, "kernels": kernels, "cpu_baseline": None, "energy_Ha": E}


def pmc_child(kind, w, steps):
    """Body of one counter pass: either a few eager solver steps or a plain
    device copy of COPY_MIB MiB, three times (the calibration kernel)."""
    import torch
    if kind == "copy":
        n = COPY_MIB * (1 << 20) // 4
        a = torch.ones(n, device="cuda", dtype=torch.float32)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        return
    if w.get("mf"):
        from realmolecule_amd.molfast import MolFast
        with MolFast(w["cfg"]) as host:
            host.do_steps(steps)
        return
    s = make_solver(w, graph=False)
    s.initialize()
    s.step(0, steps, poisson_interval=1)
    s.sync()
    s.close()


def _pmc_pass(counter, kind, workload_name, outdir, timeout_s=120):
    """Run rocprofv3 --pmc <counter> over a child bench.py process; returns
    {kernel name: [KiB per dispatch]} or None if the pass did not complete."""
    import csv
    import subprocess
    d = os.path.join(outdir, f"{kind}_{counter}")
    cmd = ["rocprofv3", "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "run", "--",
           sys.executable, os.path.join(ROOT, "bench.py"), "--pmc-child", kind, "--workload", workload_name]
    env = dict(os.environ, TMPDIR="/tmp")
    try:
        r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                           timeout=timeout_s)
    except (OSError, subprocess.TimeoutExpired) as e:
        log(f"[pmc] {counter} {kind}: {type(e).__name__}")
        return None
    path = None
    for root, _, files in os.walk(d):
        for f in files:
            if f.endswith("counter_collection.csv"):
                path = os.path.join(root, f)
    if path is None:
        log(f"[pmc] {counter} {kind}: no counter file (rc {r.returncode}): {r.stdout[-400:]!r}")
        return None
    out = {}
    with open(path) as f:
        for row in csv.DictReader(f):
            out.setdefault(row["Kernel_Name"], []).append(float(row["Counter_Value"]))
    return out


def pmc_traffic(workload_name, kernel_keys):
    """HBM bytes per dispatch of each fine-level kernel in ``kernel_keys``
    ({family: rocprofv3 name fragment}): FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch, one rocprofv3 pass each over a few eager solver steps), each
    scaled by the bytes a plain copy of known size moves over what the same
    counter reads for it.  Returns ({family: bytes}, info) or (None, why)."""
    import shutil
    import tempfile
    if shutil.which("rocprofv3") is None:
        return None, "rocprofv3 not on PATH"
    outdir = tempfile.mkdtemp(prefix="rq_pmc_", dir="/tmp")
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        for kind in ("solver", "copy"):
            log(f"[pmc] {counter} over {kind}")
            r = _pmc_pass(counter, kind, workload_name, outdir)
            if r is None:
                return None, f"{counter} pass over {kind} failed"
            res[(counter, kind)] = r
    shutil.rmtree(outdir, ignore_errors=True)
    copy_bytes = COPY_MIB * (1 << 20)
    info = {"copy_MiB": COPY_MIB}
    total = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        copies = [v for k, vs in res[(counter, "copy")].items() if "copy" in k.lower() for v in vs
                  if v * 1024 > 0.25 * copy_bytes]
        if not copies:
            return None, f"{counter}: no copy dispatches"
        scale = copy_bytes / (np.median(copies) * 1024)
        info[counter] = {"copy_KiB": round(float(np.median(copies)), 1), "scale": round(float(scale), 4)}
        for fam, keys in kernel_keys.items():
            kern = [v for k, vs in res[(counter, "solver")].items() if any(q in k for q in keys) for v in vs]
            if not kern:
                continue
            kern = [v for v in kern if v >= 0.5 * max(kern)]  # the fine level's dispatches
            per = float(np.mean(kern)) * 1024 * scale
            info[counter][fam] = {"kernel": "|".join(keys), "KiB_per_dispatch": round(float(np.mean(kern)), 1),
                                  "dispatches": len(kern), "bytes": round(per)}
            total[fam] = total.get(fam, 0.0) + per
    both = {f: v for f, v in total.items()
            if f in info["FETCH_SIZE"] and f in info["WRITE_SIZE"]}
    return (both or None), info


def cpu_baseline_c(w, fields, steps):
    """All-core CPU baseline: the C + OpenMP restatement of the same step
    (oracle/cstep.c) on the GPU's initialised fields (U, P, K, label)."""
    from oracle import cstep, s1
    NN = w["NN"]
    U, P, K, lab = fields
    g = s1.Grid(NN, w["screen"], w["dv"], np.float32)
    t0 = time.perf_counter()
    cstep.run_steps(U, P, K, lab, w["rows"][:, 3], g, w["mg_levels"], 0, steps, poisson_interval=1)
    dt = time.perf_counter() - t0
    return (NN - 1) ** 3 * steps / dt / 1e9, dt, cstep.threads()


def periodic_correction(first, K, timing, events):
    """ms to add to a K-step window starting after step ``first`` so that
    each periodic pass counts K / interval times: events = {family:
    interval}; timing = per-family {avg_us}.  Returns (ms, detail)."""
    add = 0.0
    detail = {}
    for fam, n in events.items():
        if fam not in timing:
            continue
        in_window = (first + K) // n - first // n
        per_call_ms = timing[fam]["avg_us"] * 1e-3
        corr = (K / n - in_window) * per_call_ms
        add += corr
        detail[fam] = {"interval": n, "calls_in_window": in_window, "calls_amortised": K / n,
                       "ms_per_call": round(per_call_ms, 4), "correction_ms": round(corr, 4)}
    return add, detail


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="c5",
                    choices=("c1", "c2", "c3", "c4", "c5", "mf", "pd", "sic", "hairpin", "c5forces", "c4forces",
                             "c4relax", "c3forces"))
    ap.add_argument("--force-window", type=int, default=100, help="steps of the map a full adjoint differentiates")
    ap.add_argument("--converge-max", type=int, default=10000)
    ap.add_argument("--relax-iters", type=int, default=50)
    ap.add_argument("--relax-fd-iters", type=int, default=2)
    ap.add_argument("--relax-steps", type=int, default=200)
    ap.add_argument("--cpu-steps", type=int, default=80, help="steps of the all-core C baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--timing-steps", type=int, default=200)
    ap.add_argument("--no-forces", action="store_true")
    ap.add_argument("--force-workload", default="c2", choices=("c1", "c2", "c4"))
    ap.add_argument("--force-components", type=int, default=4,
                    help="force components checked (and timed) by central differences")
    ap.add_argument("--force-batch", type=int, default=4)
    ap.add_argument("--no-pmc", action="store_true", help="skip the rocprofv3 FETCH/WRITE_SIZE passes")
    ap.add_argument("--c3-eager", action="store_true", help="c3: launch the batch's passes without graph replay")
    ap.add_argument("--c3-batch-only", action="store_true", help="c3: skip the one-handle-per-molecule comparison")
    ap.add_argument("--pmc-child", choices=("solver", "copy"), help=argparse.SUPPRESS)
    ap.add_argument("--pmc-steps", type=int, default=4)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the solver is single-GPU; "
              "multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)
    if args.pmc_child:
        pmc_child(args.pmc_child, workload(args.workload), args.pmc_steps)
        return
    if args.workload == "c3":
        print(json.dumps(bench_c3(args.steps, args.warmup, graph=not args.c3_eager, sequential=not args.c3_batch_only)))
        return
    if args.workload == "hairpin":
        print(json.dumps(bench_hairpin()))
        return
    if args.workload in ("c5forces", "c4forces"):
        print(json.dumps(bench_full_adjoint(args.workload[:2], args.force_window, args.converge_max,
                                            args.force_components, log)))
        return
    if args.workload == "c3forces":
        print(json.dumps(bench_c3_forces(args.force_window, 10, log)))
        return
    if args.workload == "c4relax":
        print(json.dumps(bench_relax("c4", args.relax_iters, args.relax_fd_iters, args.relax_steps, log)))
        return
    if args.workload == "mf":
        print(json.dumps(bench_mf(args)))
        return
    if args.workload in ("pd", "sic"):
        print(json.dumps(bench_pother(args, args.workload)))
        return

    import ctypes as C
    from realmolecule_amd import _lib as L

    w = workload(args.workload)
    NN = w["NN"]
    n3 = (NN - 1) ** 3
    log(f"[bench] workload {w['desc']}; creating solver")
    s = make_solver(w)
    w["dv"] = s.dv
    s.initialize()
    fields = None
    if not args.no_cpu:
        # the CPU baseline starts from the same initialised fields
        S = NN + 1
        fields = (s.U.get().reshape(S, S, S), s.P.get().reshape(S, S, S), s.K_buf.get().reshape(S, S, S),
                  s.label.get().reshape(S, S, S))
    sched_pi = 1
    log(f"[bench] warmup {args.warmup} steps")
    s.step(0, args.warmup, poisson_interval=sched_pi)
    s.sync()

    log(f"[bench] timed {args.steps} steps")
    t0 = time.perf_counter()
    s.step(args.warmup, args.steps, poisson_interval=sched_pi)
    s.sync()
    dt = time.perf_counter() - t0
    ms_window = dt / args.steps * 1e3
    # energy at the next step that is a multiple of 100 (normalise + boundary
    # rule just applied: where cloud/solver.py:594-595 reports it), untimed
    done = args.warmup + args.steps
    extra = (-done) % 100
    if extra:
        s.step(done, extra, poisson_interval=sched_pi)
    s._compute_energy()

    # Same steps launched eagerly (no graph replay), for comparison.
    lib = s._lib
    s._chk(lib.rq_set_graph(s._h, 0))
    first = args.warmup + args.steps + extra
    s.sync()
    t0 = time.perf_counter()
    s.step(first, args.steps, poisson_interval=sched_pi)
    s.sync()
    ms_eager = (time.perf_counter() - t0) / args.steps * 1e3
    log(f"[bench] eager launches: {ms_eager:.4f} ms/step")
    first += args.steps

    # Per-kernel device time: HIP events bracketing each launch on the
    # solver's stream, over a window that is a whole number of boundary
    # intervals (so the score rule and the normalisation are sampled)
    s._chk(lib.rq_timing_enable(s._h, 1))
    tsteps = max(100, (args.timing_steps // 100) * 100)
    s.step(first, tsteps, poisson_interval=sched_pi)
    nf = len(L.FAMILIES)
    ms = (C.c_double * nf)()
    cnt = (C.c_int64 * nf)()
    s._chk(lib.rq_timing_read(s._h, ms, cnt, nf))
    s._chk(lib.rq_timing_enable(s._h, 0))
    kernels = {}
    for f, name in enumerate(L.FAMILIES):
        if cnt[f] == 0:
            continue
        kernels[name] = {"launches": int(cnt[f]), "avg_us": round(ms[f] / cnt[f] * 1e3, 3),
                         "total_ms": round(ms[f], 4), "launches_per_step": cnt[f] / tsteps}
    # algorithmic bytes per launch: per-voxel figure x units per launch.  The
    # fine smoother family runs either one sweep per launch or a sweep pair
    # (temporal blocking on grids beyond the MALL); the fine level does
    # nu1 + nu2 = 4 sweeps per step, one of them inside smooth_itp and two
    # inside smooth_restrict or post_itp when those fused passes run, so
    # sweeps per launch = (4 - smooth_itp/step - 2 smooth_restrict/step
    # - 2 post_itp/step) / smooth/step.
    bpv = dict(BYTES_PER_VOXEL)
    if "smooth" in kernels:
        sweeps = (4 - kernels.get("smooth_itp", {}).get("launches_per_step", 0.0)
                  - 2 * kernels.get("smooth_restrict", {}).get("launches_per_step", 0.0)
                  - 2 * kernels.get("post_itp", {}).get("launches_per_step", 0.0))
        per_launch = sweeps / kernels["smooth"]["launches_per_step"]
        kernels["smooth"]["sweeps_per_launch"] = round(per_launch, 3)
        bpv["smooth"] = BYTES_PER_VOXEL["smooth"] * per_launch
    for name, ent in kernels.items():
        if name in bpv:
            ent["alg_bytes_per_voxel"] = bpv[name]
            ent["alg_bytes_per_launch"] = bpv[name] * n3
            ent["achieved_GBs"] = round(bpv[name] * n3 / (ent["avg_us"] * 1e-6) / 1e9, 1)
    # the score rule and the normalisation every 100 / 20 steps, amortised
    # into the timed window whatever its length
    corr_ms, corr = periodic_correction(args.warmup, args.steps, kernels, {"boundary": 100, "normalize": 20})
    ms_per_step = ms_window + corr_ms / args.steps
    value = n3 / (ms_per_step * 1e-3) / 1e9
    log(f"[bench] {ms_window:.4f} ms/step in the window, {ms_per_step:.4f} with the periodic passes amortised: "
        f"{value:.2f} Gvoxel-updates/s, E={s.E:.6f}")
    fine = {k: v for k, v in kernels.items() if k in bpv}
    dom = max(fine, key=lambda k: fine[k]["total_ms"])
    step_bytes = sum(bpv[k] * n3 * kernels[k]["launches"] for k in fine) / tsteps
    roof = {
        "kernel": dom,
        "bound": "hbm",
        "achieved": fine[dom]["achieved_GBs"],
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(fine[dom]["achieved_GBs"] / HBM_PEAK_GBS, 4),
        "traffic": None,
        "traffic_unit": "bytes per launch",
        "alg_bytes_per_voxel": bpv[dom],
        "alg_bytes_per_launch": bpv[dom] * n3,
        "avg_us": fine[dom]["avg_us"],
        "step_alg_bytes_fine": step_bytes,
        "step_frac": round(step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
    }
    s.close()
    if not args.no_pmc:
        keys = {k: PMC_KERNEL[k] for k in fine if k in PMC_KERNEL}
        log(f"[bench] HBM traffic of the fine kernels {sorted(keys)}: rocprofv3 FETCH_SIZE / WRITE_SIZE passes")
        traffic, info = pmc_traffic(args.workload, keys)
        roof["traffic_pmc"] = info
        if traffic:
            for k, tb in traffic.items():
                kernels[k]["traffic_bytes_per_launch"] = round(tb)
                kernels[k]["traffic_over_alg"] = round(tb / (bpv[k] * n3), 4)
                kernels[k]["hbm_GBs"] = round(tb / (kernels[k]["avg_us"] * 1e-6) / 1e9, 1)
            if dom in traffic:
                roof["traffic"] = round(traffic[dom])
                roof["traffic_over_alg"] = kernels[dom]["traffic_over_alg"]
                roof["hbm_frac"] = round(kernels[dom]["hbm_GBs"] / HBM_PEAK_GBS, 4)

    forces = None
    if not args.no_forces:
        wf = workload(args.force_workload)
        log(f"[bench] forces ({args.force_workload}): adjoint vs central differences at converged energy")
        forces = forces_bench(wf, 20, args.force_batch, 1e-3, args.force_components, log)

    cpu = None
    if fields is not None:
        log(f"[bench] cpu baseline: {args.cpu_steps} steps of the C + OpenMP restatement on the same fields")
        gv, cdt, thr = cpu_baseline_c(w, fields, args.cpu_steps)
        cpu = {"value": round(gv, 6), "unit": "Gvoxel-updates/s", "cores": thr, "kind": "port",
               "sample": f"{args.cpu_steps} steps (V-cycle + ITP, normalise every 20) of the same {NN}^3 "
                         f"workload from the GPU's initialised fields; C + OpenMP restatement (oracle/cstep.c), "
                         f"{thr} threads, {cdt:.1f} s; host cpu_count={os.cpu_count()}"}
        log(f"[bench] cpu {gv:.5f} Gvoxel-updates/s on {thr} threads")

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gvoxel-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "ms_per_step_window": round(ms_window, 5),
        "periodic_amortisation": corr,
        "ms_per_step_eager": round(ms_eager, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": w["desc"], "NN": NN, "interior_voxels": n3, "domains": len(w["rows"]),
                   "mg_levels": w["mg_levels"], "nu1": 2, "nu2": 2, "norm_interval": 20,
                   "boundary_interval": 100, "parallelism": "single-gpu"},
        "roofline": roof,
        "kernels": kernels,
        "cpu_baseline": cpu,
        "forces": forces,
        "energy_Ha": s.E,
        "energy_step": done + extra,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
