"""The batched ensemble workload (256 droplet/bubble cases, mesh side from
{64, 96, 128, 192, 256}, density ratio log-uniform in [2, 1000], surface
tension log-uniform in [1e-4, 1e-2], FP32-bulk) checked for accuracy, not only
against itself:

* cases at the corners of its parameter range -- density ratio 2, 30, 1000 x
  sigma 1e-4, 1e-2, droplets and bubbles -- batched in FP32-bulk against the
  FP64 fused path (bit-exact with the oracle, tests/test_gpu_fused_path.py)
  after 500 steps: phi, rho and p within 1e-5 relative L2, or within twice the
  distance of FP64 runs started from states perturbed at FP32 round-off (the
  problem's own sensitivity; the largest of three perturbation seeds, which
  alone range 4.2e-6..1.05e-5 on the worst case), velocity within 1e-3 of
  max(|u|, 1e-6) or twice that perturbed distance (measured: at most 1.01e-5,
  density ratio 30, sigma 1e-2);
* the full 256-case batch for 2,000 steps: every case finite with positive
  density, and sum(phi) of every case conserved to 1e-11 relative (phi is held
  in FP64 and moved only by face mass fluxes).
"""
import numpy as np
import pytest

from mpdugks_amd import abi, cases, solver

pytestmark = pytest.mark.gpu

STEPS = 500


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _corner_specs():
    specs = []
    for ratio in (2.0, 30.0, 1000.0):
        for sigma in (1e-4, 1e-2):
            specs.append({"side": 64, "radius": 16.0, "ratio": ratio, "sigma": sigma,
                          "sign": -1 if len(specs) % 2 else 1})
    return specs


def _fp64_run(spec, perturb=None):
    cfg, st = cases.ensemble_case(spec, precision=abi.MPD_FP64)
    if perturb is not None:
        rng = np.random.default_rng(perturb)
        st = st.copy()
        st.h = st.h * (1 + 6e-8 * rng.uniform(-1, 1, st.h.shape))
        st.f = st.f * (1 + 6e-8 * rng.uniform(-1, 1, st.f.shape))
    s = solver.Solver(cfg, path=abi.MPD_PATH_FUSED)
    s.upload(st)
    s.step(STEPS)
    out = s.download()
    s.close()
    return out


def test_ensemble_corners_fp32_batch_vs_fp64():
    specs = _corner_specs()
    b = solver.Batch()
    members = []
    for sp in specs:
        cfg, st = cases.ensemble_case(sp)
        s = solver.Solver(cfg, path=abi.MPD_PATH_FUSED)
        s.upload(st)
        s.step(2)
        b.add(s)
        members.append(s)
    b.step(STEPS - 2)
    bad = []
    for sp, s in zip(specs, members):
        g = s.download()
        ref = _fp64_run(sp)
        pers = [_fp64_run(sp, perturb=seed) for seed in (1, 7, 11)]
        assert np.isfinite(g.m).all()
        for k in ("phi", "rho", "p"):
            e32 = _rel(g.m[abi.M[k]], ref.m[abi.M[k]])
            floor = max(_rel(per.m[abi.M[k]], ref.m[abi.M[k]]) for per in pers)
            if e32 > max(1e-5, 2 * floor):
                bad.append((sp["ratio"], sp["sigma"], k, e32, floor))
        for k in ("u", "v"):
            du = np.abs(g.m[abi.M[k]] - ref.m[abi.M[k]]).max()
            dper = max(np.abs(per.m[abi.M[k]] - ref.m[abi.M[k]]).max() for per in pers)
            umax = max(np.abs(ref.m[abi.M["u"]]).max(), np.abs(ref.m[abi.M["v"]]).max(), 1e-6)
            if du > max(1e-3 * umax, 2 * dper):
                bad.append((sp["ratio"], sp["sigma"], k, du, dper))
    b.close()
    assert bad == []


def test_full_ensemble_finite_and_conservative():
    specs = cases.ensemble_specs(256, seed=0)
    b = solver.Batch()
    members, phi0 = [], []
    for sp in specs:
        cfg, st = cases.ensemble_case(sp)
        s = solver.Solver(cfg, path=abi.MPD_PATH_FUSED)
        s.upload(st)
        s.step(2)
        phi0.append(s.reductions()["sum_phi"])
        b.add(s)
        members.append(s)
    assert b.size() == (256, sum(sp["side"] ** 2 for sp in specs))
    b.step(2000)
    worst = 0.0
    for s, p0 in zip(members, phi0):
        r = s.reductions()
        assert r["status"] == 0 and r["nonfinite"] == 0
        worst = max(worst, abs(r["sum_phi"] - p0) / p0)
    b.close()
    assert worst < 1e-11, worst
