"""Static droplet to 10^4 steps: Laplace pressure and spurious currents.

128 x 128 periodic, AC + HCZ D2Q9, density ratio 10, sigma 1e-3, interface
width 4, radius 32 (the static-droplet workload).  The golden record holds the
FP64 oracle's state after 10^3 and 10^4 steps (tests/golden/make_droplet_golden.py).

* The FP64 fused path reproduces the oracle bit for bit over the whole run
  (plane checksums).
* The reference algorithm's own Laplace jump is Delta p = 1.88 sigma/R at
  10^4 steps (1.79 sigma/R at 5 10^4): the pressure of the HCZ closure is not
  the thermodynamic one and the droplet is still ringing at 10^4 steps; the
  record pins the value, not Laplace's law.
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


def test_fp32_bulk_fields_vs_fp64(runs):
    def rel(a, b):
        return float(np.sqrt(((a - b) ** 2).sum()) / np.sqrt((b ** 2).sum()))
    for k in (1000, 10000):
        a, b = runs[(abi.MPD_FP32_BULK, k)], runs[(abi.MPD_FP64, k)]
        floor = gold()[str(k)]["perturbed_rel_l2"]
        for name in ("rho", "phi", "p"):
            e = rel(a[abi.M[name]], b[abi.M[name]])
            bound = 1e-5 if k == 1000 else max(1e-5, 2 * floor[name])
            print("fp32 droplet step %d %s: rel L2 %.3e (bound %.3e, perturbed FP64 %.3e)" % (k, name, e, bound, floor[name]))
            assert e < bound, (k, name, e, bound)
