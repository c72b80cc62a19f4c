"""Rayleigh-Taylor 256x1024 (the benchmark workload) to its full 2 10^4 steps:
FP32-bulk against FP64 on the physical observables.

The instability amplifies round-off (the interface starts from seeded noise of
1e-3 W), so pointwise fields of two runs with different rounding drift apart;
the test bounds that drift by the problem's own sensitivity (the FP64 run from
a start perturbed at FP32 round-off size) and checks what the flow does:
* spike (lowest row holding heavy fluid) and bubble (highest row holding light
  fluid) within one cell of FP64 at 5 10^3, 10^4 and 2 10^4 steps;
* kinetic energy within 1e-3 relative;
* sum(phi) within 1e-10 relative (conserved by the scheme; the FP64 path's own
  rounding moves it by ~1e-12 over the run, FP32-bulk carries phi in FP64);
* relative L2 of phi within 2x the FP64-perturbed distance (the largest over
  three perturbation seeds), or 1e-5.
"""
import numpy as np
import pytest

from mpdugks_amd import abi, cases, solver

pytestmark = pytest.mark.gpu

STEPS = (5000, 10000, 20000)


def _run(precision, perturb=None):
    cfg, st = cases.rayleigh_taylor(nx=256, ny=1024, precision=precision)
    if perturb is not None:
        rng = np.random.default_rng(perturb)
        st.h = st.h * (1 + 6e-8 * rng.uniform(-1, 1, st.h.shape))
        st.f = st.f * (1 + 6e-8 * rng.uniform(-1, 1, st.f.shape))
    s = solver.Solver(cfg, path=abi.MPD_PATH_FUSED)
    s.upload(st)
    out, done = {}, 0
    for k in STEPS:
        s.step(k - done)
        done = k
        phi = s.download().m[abi.M["phi"]]
        r = s.reductions()
        heavy = phi.reshape(1024, 256) > 0.5
        rows = np.arange(1024)
        out[k] = {"phi": phi, "spike": int(rows[heavy.any(axis=1)].min()),
                  "bubble": int(rows[(~heavy).any(axis=1)].max()), "ke": r["kinetic_energy"],
                  "sum_phi": r["sum_phi"], "status": r["status"]}
    s.close()
    return out


def _rel(a, b):
    return float(np.sqrt(((a - b) ** 2).sum()) / np.sqrt((b ** 2).sum()))


@pytest.fixture(scope="module")
def runs():
    out = {"fp64": _run(abi.MPD_FP64), "fp32": _run(abi.MPD_FP32_BULK)}
    for seed in (1, 2, 3):
        out["fp64_perturbed_%d" % seed] = _run(abi.MPD_FP64, perturb=seed)
    return out


@pytest.mark.parametrize("k", STEPS)
def test_rt_observables_fp32_vs_fp64(runs, k):
    a, b = runs["fp32"][k], runs["fp64"][k]
    e = _rel(a["phi"], b["phi"])
    floor = max(_rel(runs["fp64_perturbed_%d" % seed][k]["phi"], b["phi"]) for seed in (1, 2, 3))
    print("rt step %d: spike %d/%d bubble %d/%d KE rel %.2e sum_phi rel %.2e phi L2 %.2e (perturbed FP64 %.2e)"
          % (k, a["spike"], b["spike"], a["bubble"], b["bubble"], abs(a["ke"] / b["ke"] - 1),
             abs(a["sum_phi"] / b["sum_phi"] - 1), e, floor))
    assert a["status"] == 0 and b["status"] == 0
    assert abs(a["spike"] - b["spike"]) <= 1 and abs(a["bubble"] - b["bubble"]) <= 1
    assert abs(a["ke"] / b["ke"] - 1) < 1e-3
    assert abs(a["sum_phi"] / b["sum_phi"] - 1) < 1e-10
    assert e < max(2 * floor, 1e-5), (e, floor)


def test_rt_mixing_layer_grows(runs):
    """The heavy fluid falls and the light rises: the mixing width grows."""
    w = [runs["fp64"][k]["bubble"] - runs["fp64"][k]["spike"] for k in STEPS]
    assert w[0] < w[1] < w[2] and w[2] > 200
