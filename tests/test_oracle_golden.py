"""The CPU oracle against the reference's own results (golden digests).

tests/golden/ref_*.json hold SHA-256 digests of the full cell state of the
compiled reference (Ark, unmodified sources) after 0, 1, 2, 10 and 100 steps of
its AC_Drop case on its 128x128 mesh (D2Q9HCZ, D2Q9AC, and both with walls in
y: bounce-back for HCZ, non-equilibrium extrapolation for HCZ and AC-p).  The host initialiser and the oracle must reproduce
every digest bit for bit.
"""
import json
import os

import numpy as np
import pytest

from mpdugks_amd import abi, solver
from helpers import digest, make_case, oracle_run

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
CASES = ["hcz", "acp", "hcz_wall", "hcz_nee", "acp_nee"]
UW_CASES = ["hcz_uw", "acp_uw"]


def load(name):
    with open(os.path.join(GOLDEN, "ref_%s.json" % name)) as fh:
        return json.load(fh)


def check_state(st, rec):
    bad = []
    if "h" in rec and digest(st.h) != rec["h"]:
        bad.append("h")
    if digest(st.f) != rec["f"]:
        bad.append("f")
    for k, v in rec["m"].items():
        if digest(st.m[abi.M[k]]) != v:
            bad.append(k)
    return bad


@pytest.mark.parametrize("name", CASES + UW_CASES + ["pp"])
def test_initial_state_matches_reference(name):
    g = load(name)
    cfg, st = make_case(g["model"], walls=g["walls_y"], nee=g.get("bc") == "nee")
    assert check_state(st, g["steps"]["0"]) == []


@pytest.mark.parametrize("name", UW_CASES)
def test_oracle_uw_first_step_bit_exact(name, oracle_mod):
    g = load(name)
    cfg, st = make_case(g["model"], flux=abi.MPD_FLUX_UW)
    out = oracle_run(cfg, st, [1])
    assert check_state(out[1], g["steps"]["1"]) == []


@pytest.mark.parametrize("name", CASES)
def test_oracle_bit_exact_vs_reference(name, oracle_mod):
    g = load(name)
    cfg, st = make_case(g["model"], walls=g["walls_y"], nee=g.get("bc") == "nee")
    steps = sorted(int(k) for k in g["steps"] if int(k) > 0)
    out = oracle_run(cfg, st, steps)
    for k in steps:
        rec = g["steps"][str(k)]
        bad = check_state(out[k], rec)
        sample = np.array(out[k].m[abi.M["phi"]].reshape(128, 128)[64, 10:30])
        assert bad == [], "step %d mismatching planes %s (phi sample diff %.3e)" % (
            k, bad, np.abs(sample - np.array(rec["sample"]["phi"])).max())


def test_golden_detects_a_planted_mistake(oracle_mod):
    """A 1e-12 relative change of one population in the start state must change
    the step-10 digests: the digests pin the state far below any tolerance."""
    g = load("hcz")
    cfg, st = make_case(g["model"])
    st.f[3, 64 * 128 + 20] *= 1.0 + 1e-12
    out = oracle_run(cfg, st, [10])
    assert check_state(out[10], g["steps"]["10"]) != []
