"""Conditioning of the reference algorithm (CPU, oracle only).

The conservative Allen-Cahn source of the reference normalises grad(phi)
(Ark/Cpp/SourceM.cpp:47-58); where phi is not exactly 0 or 1 but its gradient
is near zero the direction is set by round-off, so any change in rounding is
amplified.  This pins the scale: a 1e-12 relative perturbation of the FP64
state of the 96x96 HCZ bubble grows past 1e-6 relative L2 in phi by step 200.
Any implementation with different rounding (FP32-bulk in particular) is
therefore compared with the FP64 oracle over horizons shorter than this.
"""
import numpy as np

from mpdugks_amd import abi
from helpers import make_case, oracle_run


def rel_l2(a, b):
    return float(np.sqrt(((a - b) ** 2).sum()) / np.sqrt((b ** 2).sum()))


def test_hcz_bubble_amplifies_round_off(oracle_mod):
    cfg, st = make_case("ac_hcz", nx=96, ny=96, radius=30.0)
    ref = oracle_run(cfg, st, [60, 200])
    st2 = st.copy()
    rng = np.random.default_rng(1)
    st2.h *= 1 + 1e-12 * rng.uniform(-1, 1, st2.h.shape)
    st2.f *= 1 + 1e-12 * rng.uniform(-1, 1, st2.f.shape)
    per = oracle_run(cfg, st2, [60, 200])
    e60 = rel_l2(per[60].m[abi.M["phi"]], ref[60].m[abi.M["phi"]])
    e200 = rel_l2(per[200].m[abi.M["phi"]], ref[200].m[abi.M["phi"]])
    assert e60 < 1e-11          # still at the perturbation's scale at 60 steps
    assert e200 > 1e-6          # amplified by > 1e6 by 200 steps
