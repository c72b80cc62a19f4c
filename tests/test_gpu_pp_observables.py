"""Pseudopotential D2Q9 (Carnahan-Starling EOS, T/Tc = 0.85), 512x512 periodic:
FP32-bulk against the plain FP64 path (bit-exact with the oracle) on what the
flow produces, at the horizons the reference algorithm survives.

* The reference's own start (SCMP_Drop, Ark/Cpp/Preprocess.cpp:337, scaled to
  512^2) for 5 10^4 steps: stable in both precisions.
* The noisy spinodal start (rho0 = 0.1645 (1 +- 1%) per cell) for 2,000 steps.
  The reference algorithm itself does not survive this start much longer: the
  FP64 path (and FP32-bulk alike) is non-finite everywhere by step 3,000 (it
  is finite at 2,500), so 5 10^4 steps of it are not a meaningful target.

Observables (pointwise fields of the two precisions drift apart under the
phase-separation dynamics): the vapour and liquid densities (1st and 99th
percentiles of rho) within 2e-3 relative, the liquid fraction (cells above
the mean of the two) within 0.005, FP32-bulk mass conserved to 1e-12 (rho is
carried in FP64), every cell finite.
"""
import numpy as np
import pytest

from mpdugks_amd import abi, solver
from mpdugks_amd.config import default_config
from helpers import make_case

pytestmark = pytest.mark.gpu


def _run(start, steps, precision):
    if start == "drop":
        cfg, st = make_case("pp_d2q9", nx=512, ny=512, precision=precision)
        s = solver.Solver(cfg)
        s.upload(st)
    else:
        cfg = default_config(abi.MPD_MODEL_PP_D2Q9, nx=512, ny=512, pp_tr=0.85)
        cfg.precision = precision
        s = solver.Solver(cfg)
        s.init_pp_noise(0.1645, amp=0.01, seed=3)
    m0 = s.reductions()["sum_rho"]
    s.step(steps)
    r = s.reductions()
    rho = s.download().m[abi.M["rho"]]
    s.close()
    return rho, r, m0


@pytest.mark.parametrize("start,steps", [("drop", 50000), ("noise", 2000)])
def test_pp_coexistence_fp32_vs_fp64(start, steps):
    (a, ra, ma), (b, rb, mb) = _run(start, steps, abi.MPD_FP32_BULK), _run(start, steps, abi.MPD_FP64)
    assert ra["status"] == 0 and rb["status"] == 0
    assert np.isfinite(a).all() and np.isfinite(b).all()
    va, la = np.percentile(a, [1, 99])
    vb, lb = np.percentile(b, [1, 99])
    mid = 0.5 * (vb + lb)
    fa, fb = float((a > mid).mean()), float((b > mid).mean())
    print("pp 512^2 %s %d steps: vapour %.6f / %.6f, liquid %.6f / %.6f, liquid fraction %.4f / %.4f "
          "(FP32 / FP64); mass drift FP32 %.2e, FP64 %.2e"
          % (start, steps, va, vb, la, lb, fa, fb, abs(ra["sum_rho"] / ma - 1), abs(rb["sum_rho"] / mb - 1)))
    assert lb > 2 * vb, "no phase separation"
    assert abs(va / vb - 1) < 2e-3 and abs(la / lb - 1) < 2e-3
    assert abs(fa - fb) < 0.005
    assert abs(ra["sum_rho"] / ma - 1) < 1e-12
