"""Static droplet to 10^4 steps: Laplace pressure and spurious currents.

128 x 128 periodic, AC + HCZ D2Q9, density ratio 10, sigma 1e-3, interface
width 4, radius 32 (the static-droplet workload).  The golden record holds the
FP64 oracle's state after 10^3 and 10^4 steps (tests/golden/make_droplet_golden.py).

* The FP64 fused path reproduces the oracle bit for bit over the whole run
  (plane checksums).
* The reference algorithm's own Laplace jump is Delta p = 1.88 sigma/R at
  10^4 steps (1.79 sigma/R at 5 10^4): the pressure of the HCZ closure is not
  the thermodynamic one and the droplet is still ringing at 10^4 steps; the
  record pins the value.  Laplace's law itself (Delta p R constant over radii,
  thermoPressure included) is checked after 5e4 steps at the end of this file.
* FP32-bulk (stated tolerances): Delta p within 1e-3 and the largest spurious
  speed within 1e-2 of the FP64 values; sum(phi) within 1e-9 (the phase field
  is carried in FP64 and moved only by antisymmetric face fluxes).  Pointwise,
  rho, phi and p within 1e-5 relative L2 after 10^3 steps; after 10^4 steps
  the problem's own conditioning sets the floor: the FP64 oracle started from
  a state perturbed at FP32 round-off size (relative 6e-8) ends 2.3e-5 (rho)
  away from itself, and FP32-bulk must stay within twice that distance
  (or 1e-5, whichever is larger).
"""
import hashlib
import json
import os

import numpy as np
import pytest

from mpdugks_amd import abi, solver
from mpdugks_amd.config import default_config

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def gold():
    with open(os.path.join(HERE, "golden", "oracle_droplet_128.json")) as fh:
        return json.load(fh)


def droplet(precision):
    cfg = default_config(abi.MPD_MODEL_AC_HCZ, nx=128, ny=128, sigma=1e-3, width=4.0)
    cfg.precision = precision
    st = solver.ac_drop(cfg, cx=64.0, cy=64.0, radius=32.0, sign=-1)
    s = solver.Solver(cfg, path=abi.MPD_PATH_FUSED)
    s.upload(st)
    return s


def observables(m):
    x = np.arange(128) + 0.5
    X, Y = np.meshgrid(x, x)
    r = np.hypot(X - 64, Y - 64).ravel()
    p, u, v = m[abi.M["p"]], m[abi.M["u"]], m[abi.M["v"]]
    return {"dp": float(p[r < 24].mean() - p[r > 40].mean()), "umax": float(np.hypot(u, v).max()),
            "sum_phi": float(m[abi.M["phi"]].sum())}


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


@pytest.fixture(scope="module")
def runs():
    out = {}
    for prec in (abi.MPD_FP64, abi.MPD_FP32_BULK):
        s = droplet(prec)
        done = 0
        for k in (1000, 10000):
            s.step(k - done)
            done = k
            out[(prec, k)] = s.download().m
        s.close()
    return out


@pytest.mark.parametrize("k", [1000, 10000])
def test_fp64_fused_bit_exact_over_the_run(runs, k):
    g = gold()[str(k)]
    m = runs[(abi.MPD_FP64, k)]
    for name in ("rho", "phi", "u", "v", "p"):
        assert sha(m[abi.M[name]]) == g["sha256_" + name], name
    assert observables(m)["dp"] == g["dp"]


def test_laplace_jump_of_the_reference_algorithm():
    g = gold()["10000"]
    sigma, R = 1e-3, 32.0
    assert 1.8 < g["dp"] / (sigma / R) < 1.95
    assert g["umax"] < 1e-5


@pytest.mark.parametrize("k", [1000, 10000])
def test_fp32_bulk_laplace_and_spurious_currents(runs, k):
    g = gold()[str(k)]
    o = observables(runs[(abi.MPD_FP32_BULK, k)])
    print("fp32 droplet step %d: dp %.9e (fp64 %.9e) umax %.6e (fp64 %.6e) sum_phi %.12e (fp64 %.12e)"
          % (k, o["dp"], g["dp"], o["umax"], g["umax"], o["sum_phi"], g["sum_phi"]))
    assert abs(o["dp"] / g["dp"] - 1) < 1e-3
    assert abs(o["umax"] / g["umax"] - 1) < 1e-2
    assert abs(o["sum_phi"] / g["sum_phi"] - 1) < 1e-9


def _perturbed_fp64(seed, steps):
    """The FP64 fused run (bit-exact with the oracle) from the start state with
    every population multiplied by 1 + 6e-8 U(-1, 1): the problem's own
    sensitivity to a round-off-sized difference."""
    cfg = default_config(abi.MPD_MODEL_AC_HCZ, nx=128, ny=128, sigma=1e-3, width=4.0)
    st = solver.ac_drop(cfg, cx=64.0, cy=64.0, radius=32.0, sign=-1)
    rng = np.random.default_rng(seed)
    st.h = st.h * (1 + 6e-8 * rng.uniform(-1, 1, st.h.shape))
    st.f = st.f * (1 + 6e-8 * rng.uniform(-1, 1, st.f.shape))
    s = solver.Solver(cfg, path=abi.MPD_PATH_FUSED)
    s.upload(st)
    out, done = {}, 0
    for k in steps:
        s.step(k - done)
        done = k
        out[k] = s.download().m
    s.close()
    return out


def test_fp32_bulk_fields_vs_fp64(runs):
    """rho, phi, p of FP32-bulk against FP64: 1e-5 relative L2 after 10^3 steps;
    after 10^4, twice the largest distance of three perturbed FP64 runs (seeds
    1, 2, 3; the golden record holds seed 1's, computed by the oracle), or 1e-5."""
    def rel(a, b):
        return float(np.sqrt(((a - b) ** 2).sum()) / np.sqrt((b ** 2).sum()))
    pert = [_perturbed_fp64(seed, (1000, 10000)) for seed in (1, 2, 3)]
    for k in (1000, 10000):
        a, b = runs[(abi.MPD_FP32_BULK, k)], runs[(abi.MPD_FP64, k)]
        for name in ("rho", "phi", "p"):
            e = rel(a[abi.M[name]], b[abi.M[name]])
            floors = [rel(q[k][abi.M[name]], b[abi.M[name]]) for q in pert]
            if k == 10000:   # seed 1 on the GPU equals the oracle's golden value
                assert abs(floors[0] / gold()[str(k)]["perturbed_rel_l2"][name] - 1) < 1e-6
            bound = 1e-5 if k == 1000 else max(1e-5, 2 * max(floors))
            print("fp32 droplet step %d %s: rel L2 %.3e (bound %.3e, perturbed FP64 %s)"
                  % (k, name, e, bound, " ".join("%.3e" % f for f in floors)))
            assert e < bound, (k, name, e, bound)


# ---------------- Laplace's law and spurious currents ----------------

def _midx_tp(tmp_path, cfg, m):
    """The thermodynamic pressure along the column i = nx/2 - 1 as the
    reference's own MidX profile writer prints it (thermoPressure,
    Ark/Cpp/Output.cpp:146-163): (y, TP)."""
    path = str(tmp_path / "midx.dat")
    solver.write_profile(path, abi.MPD_OUT_MIDX, cfg, solver.State(cfg, None, None, m), solver.output_info(cfg))
    lines = open(path).read().split("\n")
    k = next(i for i, ln in enumerate(lines) if ln.startswith("variables = Y,TP"))
    rows = [ln.split() for ln in lines[k + 2:k + 2 + cfg.ny]]
    return np.array([float(a) for a, _ in rows]), np.array([float(b) for _, b in rows])


def _laplace_run(tmp_path, radius, sigma, steps=50000):
    cfg = default_config(abi.MPD_MODEL_AC_HCZ, nx=128, ny=128, sigma=sigma, width=4.0)
    st = solver.ac_drop(cfg, cx=64.0, cy=64.0, radius=radius, sign=-1)
    s = solver.Solver(cfg, path=abi.MPD_PATH_FUSED)
    s.upload(st)
    s.step(steps)
    m = s.download().m
    s.close()
    x = np.arange(128) + 0.5
    X, Y = np.meshgrid(x, x)
    r = np.hypot(X - 64, Y - 64).ravel()
    W = cfg.width
    p = m[abi.M["p"]]
    dp = p[r < radius - 2 * W].mean() - p[r > radius + 2 * W].mean()
    y, tp = _midx_tp(tmp_path, cfg, m)
    d = np.abs(y - 64.0)
    dtp = tp[d < radius - 2 * W].mean() - tp[d > radius + 2 * W].mean()
    return dp, dtp, float(np.hypot(m[abi.M["u"]], m[abi.M["v"]]).max())


def test_laplace_law_and_spurious_currents(tmp_path):
    """Droplets of radius 16, 24, 32, 40 (128^2, density ratio 10, interface
    width 4) after 5e4 steps, FP64 fused path (bit-exact with the oracle, which
    is bit-exact with the reference):
    * Laplace's law Delta p = sigma_eff / R: Delta p * R is the same for every
      radius within 1% (measured 1.793..1.802 sigma, a 0.53% spread, the
      smallest radius highest: an O(W/R) correction), for the hydrodynamic p
      and for the thermodynamic pressure of the reference's MidX writer, which
      agree to 0.2% (in the bulk phases thermoPressure is p);
    * sigma_eff = 1.795 sigma +- 1% at this width (1.88 sigma at width 8): a
      property of the reference algorithm's pressure closure, not of the
      implementation -- the jump is not sigma / R;
    * the largest spurious speed is linear in sigma (Stokes regime): x10 for
      sigma 1e-3 -> 1e-2 within 1% (6.20e-6 -> 6.21e-5 at R = 32)."""
    sigma = 1e-3
    res = {R: _laplace_run(tmp_path, R, sigma) for R in (16.0, 24.0, 32.0, 40.0)}
    k = np.array([dp * R / sigma for R, (dp, _, _) in res.items()])
    kt = np.array([dtp * R / sigma for R, (_, dtp, _) in res.items()])
    print("Laplace: dp R / sigma", k, "thermo", kt)
    assert k.max() / k.min() - 1 < 1e-2
    assert np.abs(kt / k - 1).max() < 2e-3
    assert abs(k.mean() / 1.795 - 1) < 1e-2
    u3 = res[32.0][2]
    u2 = _laplace_run(tmp_path, 32.0, 1e-2)[2]
    print("spurious speed R=32: sigma 1e-3 %.4e, sigma 1e-2 %.4e" % (u3, u2))
    assert 1e-7 < u3 < 1e-5
    assert abs(u2 / u3 / 10 - 1) < 1e-2
