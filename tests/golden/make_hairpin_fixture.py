"""Writes tests/golden/hairpin_a100.json: the energies and fold angle of every
snapshot, and the nucleus positions at the first and last snapshot, of the
reference's one recorded solver run (cloud/hairpin_results.json: the bent
solvated hairpin of cloud/run_hairpin.py on an A100, 10000 steps).  Reads
the JSON data file only."""
import json
import os
import sys

SRC = os.path.join(sys.argv[1] if len(sys.argv) > 1 else "/root/reference", "cloud", "hairpin_results.json")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hairpin_a100.json")

d = json.load(open(SRC))
keys = ("step", "E", "E_T", "E_eK", "E_ee", "E_KK", "E_bind", "ms_per_step", "fold_angle")
out = {"source": "cloud/hairpin_results.json (reference run on an NVIDIA A100, cloud/run_hairpin.py)",
       "snapshots": [{k: s[k] for k in keys} for s in d["snapshots"]],
       "nucPos_first": d["snapshots"][0]["nucPos"], "nucPos_last": d["snapshots"][-1]["nucPos"],
       "final_E": d["final_E"]}
json.dump(out, open(OUT, "w"), indent=0)
print(f"wrote {OUT}: {len(out['snapshots'])} snapshots")
