#!/usr/bin/env python3
"""Throughput of the RealQM grid-solver step on one MI355X.

Workload (default): 10 H2O, seeded random orientations, O-O >= 2.6 A, in a
12 A cube; 192^3 FP32 grid (h = 0.1181 bohr), 5-level multigrid V-cycle per
ITP step, per-domain normalise every 20 steps, nuclear-score boundary rule
every 100 steps, FP64 reductions.  A "step" is one ITP update plus one
V-cycle.  Inputs are resident in HBM before the timed region.

Prints ONE JSON line: value = Gvoxel-updates/s = (NN-1)^3 * steps / time,
plus ``roofline`` for the dominant kernel (algorithmic bytes per launch /
average launch time from HIP events on the solver's stream) and
``cpu_baseline`` (the NumPy FP32 restatement on a bounded sample).
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
# once) of the fine-level passes.  On grids above 65^3 nodes the solver runs
# residual + restriction and prolongation + correction as separate passes;
# the last post-sweep is always fused with the ITP update.
BYTES_PER_VOXEL = {
    "itp": 18.0,         # U, K, P (f32) + label (u16) in; U' out  (steps without a V-cycle)
    "smooth": 12.0,      # P, U in; P' out
    "restrict": 16.5,    # residual (P, U in; r out) + restriction (r in; coarse rhs out, 1/8 of 4 B)
    "prolong": 8.5,      # P in and out, coarse correction in (1/8)
    "smooth_itp": 22.0,  # last post-sweep + ITP: P, U, K, label in; P', U' out
}
BYTES_FUSED_SMALL = {"restrict": 8.5, "prolong": 12.5}  # one-pass variants on grids <= 65^3 nodes


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


def cpu_baseline(w, steps):
    """NumPy FP32 restatement of the same step (oracle), single process."""
    from oracle import s1, mg
    NN, screen, rows = w["NN"], w["screen"], w["rows"]
    st = s1.State(rows, NN, screen, np.float32)
    g = st.g
    st.U, st.label, st.K, _ = s1.init_trial(rows, g)
    st.U, _ = s1.normalize(st.U, st.label, st.Z, g)
    a, c = 1.0 / 7.0, 7.0
    t0 = time.perf_counter()
    for step in range(1, steps + 1):
        rhs = s1.rho_rhs(st.U, g.two_pi, g.h2, np.float32)
        st.P = mg.vcycle(st.P, rhs, w["mg_levels"], 2, 2, a, c, np.float32)
        st.U = s1.itp_update(st.U, st.K, st.P, st.label, g)
        if step % 20 == 0:
            st.U, _ = s1.normalize(st.U, st.label, st.Z, g)
    dt = time.perf_counter() - t0
    n3 = (NN - 1) ** 3
    return n3 * steps / dt / 1e9, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--timing-steps", type=int, default=100)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the solver is single-GPU; "
              "multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    import ctypes as C
    from realmolecule_amd import _lib as L

    w = workload(args.workload)
    NN = w["NN"]
    n3 = (NN - 1) ** 3
    log(f"[bench] workload {w['desc']}; creating solver")
    s = make_solver(w)
    s.initialize()
    sched_pi = 1
    log(f"[bench] warmup {args.warmup} steps")
    s.step(0, args.warmup, poisson_interval=sched_pi)
    s.sync()

    log(f"[bench] timed {args.steps} steps")
    t0 = time.perf_counter()
    s.step(args.warmup, args.steps, poisson_interval=sched_pi)
    s.sync()
    dt = time.perf_counter() - t0
    ms_per_step = dt / args.steps * 1e3
    value = n3 * args.steps / dt / 1e9
    s._compute_energy()
    log(f"[bench] {ms_per_step:.4f} ms/step, {value:.2f} Gvoxel-updates/s, E={s.E:.6f}")

    # Same steps launched eagerly (no graph replay), for comparison.
    lib = s._lib
    s._chk(lib.rq_set_graph(s._h, 0))
    first = args.warmup + args.steps
    s.sync()
    t0 = time.perf_counter()
    s.step(first, args.steps, poisson_interval=sched_pi)
    s.sync()
    ms_eager = (time.perf_counter() - t0) / args.steps * 1e3
    log(f"[bench] eager launches: {ms_eager:.4f} ms/step")
    first += args.steps

    # Per-kernel device time: HIP events bracketing each launch on the solver's stream.
    s._chk(lib.rq_timing_enable(s._h, 1))
    s.step(first, args.timing_steps, poisson_interval=sched_pi)
    nf = len(L.FAMILIES)
    ms = (C.c_double * nf)()
    cnt = (C.c_int64 * nf)()
    s._chk(lib.rq_timing_read(s._h, ms, cnt, nf))
    s._chk(lib.rq_timing_enable(s._h, 0))
    bpv = dict(BYTES_PER_VOXEL)
    if NN + 1 <= 65:
        bpv.update(BYTES_FUSED_SMALL)
    kernels = {}
    for f, name in enumerate(L.FAMILIES):
        if cnt[f] == 0:
            continue
        avg_us = ms[f] / cnt[f] * 1e3
        ent = {"launches": int(cnt[f]), "avg_us": round(avg_us, 3), "total_ms": round(ms[f], 4)}
        if name in bpv:
            gbs = bpv[name] * n3 / (avg_us * 1e-6) / 1e9
            ent["alg_bytes_per_launch"] = bpv[name] * n3
            ent["achieved_GBs"] = round(gbs, 1)
        kernels[name] = ent
    fine = {k: v for k, v in kernels.items() if k in bpv}
    dom = max(fine, key=lambda k: fine[k]["total_ms"])
    step_bytes = sum(bpv[k] * n3 * kernels[k]["launches"] for k in fine) / args.timing_steps
    roof = {
        "kernel": dom,
        "bound": "hbm",
        "achieved": fine[dom]["achieved_GBs"],
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(fine[dom]["achieved_GBs"] / HBM_PEAK_GBS, 4),
        "traffic": None,
        "alg_bytes_per_voxel": bpv[dom],
        "step_alg_bytes_fine": step_bytes,
        "step_frac": round(step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
    }

    cpu = None
    if not args.no_cpu:
        log(f"[bench] cpu baseline: {args.cpu_steps} steps of the NumPy FP32 oracle")
        gv, cdt = cpu_baseline(w, args.cpu_steps)
        cpu = {"value": round(gv, 6), "unit": "Gvoxel-updates/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_steps} steps (V-cycle + ITP, normalise every 20) of the same "
                         f"{NN}^3 workload, NumPy FP32 restatement, {cdt:.1f} s; "
                         f"host cpu_count={os.cpu_count()}"}
        log(f"[bench] cpu {gv:.5f} Gvoxel-updates/s")

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gvoxel-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
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
        "energy_Ha": s.E,
    }
    print(json.dumps(out))
    s.close()


if __name__ == "__main__":
    main()
