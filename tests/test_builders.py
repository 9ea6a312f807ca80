"""Input builders against the reference's recorded inputs (CPU only)."""
import json
import os

import numpy as np

from realmolecule_amd.builders import hairpin_solvated

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hairpin_a100.json")


def test_hairpin_builder_matches_reference_positions():
    """The waters never move in the reference run: their positions in its
    first snapshot are the builder's rounded grid nodes, exactly."""
    cfg = hairpin_solvated()
    ref = json.load(open(FIX))
    pos = np.array([[a["i"], a["j"], a["k"]] for a in cfg["atoms"]], float)
    first = np.array(ref["nucPos_first"])
    n_p = cfg["nProtein"]
    assert len(cfg["atoms"]) == len(first) == 267 and n_p == 87
    assert np.array_equal(pos[n_p:], first[n_p:])
    # the protein has moved at most one clipped step pair by the first snapshot
    assert np.abs(pos[:n_p] - first[:n_p]).max() < 0.05
