"""Node.js N-API addon: exports and error handling (CPU), and the JS host's
molecule.js frame schedule against the S2 oracle schedule (GPU)."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "node", "rq_addon.node")
NODE = shutil.which("node")

needs_node = pytest.mark.skipif(NODE is None or not os.path.exists(ADDON), reason="node or rq_addon.node missing")

EXPORTS = ["abiVersion", "dvForElectrons", "create", "destroy", "info", "initTrial", "initialize", "poisson",
           "twogridCorrect", "update", "normalize", "boundary", "recomputeK", "step", "sync", "energy", "forcesHF",
           "forcesGradP", "readField", "displaySlice", "potherFrame", "writeField", "setAtoms", "setLockGroups", "adjointGradient",
           "setAtomsS2", "adjointSaveStart"]


def run_node(code, timeout=60):
    return subprocess.run([NODE, "-e", code], capture_output=True, text=True, timeout=timeout, cwd=ROOT)


@needs_node
def test_addon_exports_and_errors():
    code = """
const a = require('./node/rq_addon.node');
const out = {keys: Object.keys(a), abi: a.abiVersion(), dv: [a.dvForElectrons(3), a.dvForElectrons(150)]};
try { a.create({NN: 9, screen: 3}, [[4, 4, 4, 1, 0]]); out.err = 'none'; } catch (e) { out.err = e.message; }
try { a.create({NN: 32, screen: 3}, 'x'); } catch (e) { out.err2 = e.message; }
// typed-array types are checked before the handle: no device call is made
const te = (f) => { try { f(); return 'none'; } catch (e) { return e.constructor.name + ': ' + e.message; } };
out.wf = [te(() => a.writeField({}, 'U', new Uint8Array(64))), te(() => a.writeField({}, 'U', new Float64Array(8))),
          te(() => a.writeField({}, 'label', new Float32Array(8))), te(() => a.writeField({}, 'U', [1, 2])),
          te(() => a.setLockGroups({}, new Uint32Array(3))), te(() => a.writeField({}, 'U', new Float32Array(8)))];
console.log(JSON.stringify(out));
"""
    r = run_node(code)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert sorted(out["keys"]) == sorted(EXPORTS)
    assert out["abi"] == 1 and out["dv"] == [0.12, 0.02]
    assert "NN=9" in out["err"]
    assert "atoms must be an array" in out["err2"]
    wf = out["wf"]
    assert wf[0] == wf[1] == "TypeError: writeField: field needs a Float32Array"
    assert wf[2] == "TypeError: writeField: label needs an Int32Array"
    assert wf[3] == "TypeError: writeField: field needs a Float32Array"
    assert wf[4] == "TypeError: setLockGroups: expected an Int32Array"
    # a well-typed array reaches the handle check
    assert wf[5] == "TypeError: expected a solver handle"


@needs_node
@pytest.mark.gpu
@pytest.mark.parametrize("activate", [None, 2])
def test_js_host_frames_match_s2_oracle(tmp_path, activate):
    """Three 20-step frames of the JS host (two-grid correction every 8 steps,
    rc rule, level set, frame potential with SIC; the direct Pother turned
    off) against the S2 oracle
    schedule in FP32; with activate = 2, after activateAtoms(2) (the third
    nucleus uncharged)."""
    from oracle import s2
    from tests.helpers import three_atoms, rel_l2
    NN, screen = 24, 24 * 0.3
    rows = three_atoms(NN)
    rows[1, 4] = 0.7
    cfg = {"USER_NN": NN, "USER_SCREEN": screen,
           "USER_ATOMS": [{"i": float(r[0]), "j": float(r[1]), "k": float(r[2]), "Z": float(r[3])} for r in rows],
           "USER_Z": [float(z) for z in rows[:, 3]], "USER_RC": [float(x) for x in rows[:, 4]],
           "USER_STEPS_PER_FRAME": 20, "USER_POISSON_INTERVAL": 8, "USER_DIRECT_POTHER": False}
    spec = {"cfg": cfg, "frames": 3, "steps": 20, "activate": activate}
    print("running node host", flush=True)
    r = subprocess.run([NODE, os.path.join(ROOT, "node", "test_host_gpu.js"), str(tmp_path), json.dumps(spec)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    frames = json.loads((tmp_path / "node_frames.json").read_text())
    S = NN + 1
    U = np.fromfile(tmp_path / "node_U.f32", np.float32).reshape(S, S, S)
    L = np.fromfile(tmp_path / "node_label.i32", np.int32).reshape(S, S, S)
    # the host's default frame potential: PotherBuf with SIC every frame (3 electrons)
    orows = rows.copy()
    if activate:
        orows[activate:, 3] = 0.0
    st = s2.S2State(orows, NN, screen, np.float32, dv=0.12, poisson_interval=8, pother_frame=True, sic_interval=1)
    for f in range(3):
        tot, per, dip = st.frame(20)
        E_el = tot.sum()
        node_el = frames[f]["E_T"] + frames[f]["E_eK"] + frames[f]["E_ee"]
        print(f"frame {f}: node {node_el:.9f} oracle {E_el:.9f}")
        assert abs(node_el - E_el) <= 1e-6 * abs(E_el)
    assert rel_l2(U.astype(np.float64) ** 2, st.U.astype(np.float64) ** 2) < 1e-5
    assert (L != st.label).sum() <= 2


@needs_node
@pytest.mark.gpu
def test_js_host_direct_pother_and_atom_fields_match_s2_oracle(tmp_path):
    """The page's defaults for four electrons: the direct Pother (10 damped
    sweeps of each P_m per step), here with a bare nucleus (USER_Z_NUC) and a
    per-atom INIT_ZEFF; three 20-step frames against the S2 oracle in FP32,
    and E_KK over the unique nuclei (the first two electrons share one
    nucleus of charge 2, the bare nucleus counts with its Z_nuc)."""
    from oracle import s2
    from tests.helpers import rel_l2
    NN, screen = 24, 24 * 0.3
    c = NN / 2
    rows = np.array([[c - 3.3, c + 0.7, c + 0.2, 1.0, 0.0],
                     [c - 3.3, c + 0.7, c + 0.2, 1.0, 0.0],
                     [c + 2.4, c - 0.3, c - 0.6, 2.0, 0.4],
                     [c + 3.1, c + 4.9, c + 0.9, 0.0, 0.0]])
    Z_nuc = [1.0, 1.0, 2.0, 1.0]
    zeff = [0.8, 0.0, 0.0, 0.0]
    cfg = {"USER_NN": NN, "USER_SCREEN": screen,
           "USER_ATOMS": [{"i": float(r[0]), "j": float(r[1]), "k": float(r[2]), "Z": float(r[3])} for r in rows],
           "USER_Z": [float(z) for z in rows[:, 3]], "USER_RC": [float(x) for x in rows[:, 4]],
           "USER_Z_NUC": Z_nuc, "USER_INIT_ZEFF": zeff,
           "USER_STEPS_PER_FRAME": 20, "FREEZE_BOUNDARY": False}
    spec = {"cfg": cfg, "frames": 3, "steps": 20}
    print("running node host (direct Pother)", flush=True)
    r = subprocess.run([NODE, os.path.join(ROOT, "node", "test_host_gpu.js"), str(tmp_path), json.dumps(spec)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    frames = json.loads((tmp_path / "node_frames.json").read_text())
    S = NN + 1
    U = np.fromfile(tmp_path / "node_U.f32", np.float32).reshape(S, S, S)
    L = np.fromfile(tmp_path / "node_label.i32", np.int32).reshape(S, S, S)
    ext = [{"init_zeff": 0.8}, {}, {}, {"Z_nuc": 1.0}]
    st = s2.S2State(rows, NN, screen, np.float32, dv=0.12, direct=True, ext=ext)
    ekk = s2.nuclear_energy_unique([tuple(r[:3]) for r in rows], rows[:, 3], Z_nuc, screen / NN)
    for f in range(3):
        tot, per, dip = st.frame(20)
        E_el = tot.sum()
        node_el = frames[f]["E_T"] + frames[f]["E_eK"] + frames[f]["E_ee"]
        print(f"frame {f}: node {node_el:.9f} oracle {E_el:.9f}; E_KK {frames[f]['E_KK']:.12f} {ekk:.12f}")
        assert abs(node_el - E_el) <= 1e-6 * abs(E_el)
        assert frames[f]["E_KK"] == pytest.approx(ekk, rel=1e-14)
    assert rel_l2(U.astype(np.float64) ** 2, st.U.astype(np.float64) ** 2) < 1e-5
    assert (L != st.label).sum() <= 2
    assert (L == 3).sum() == 0   # the bare nucleus owns no node
