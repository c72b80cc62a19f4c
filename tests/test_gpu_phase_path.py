"""GPU plain FP64 path (one HIP kernel per reference phase) against the oracle.

Compiled without floating-point contraction, the phase kernels evaluate the
reference's expressions in the reference's order, so the GPU state must equal
the oracle's bit for bit; the oracle itself is bit-exact with the reference
(test_oracle_golden.py).
"""
import numpy as np
import pytest

from mpdugks_amd import abi, solver
from helpers import make_case, max_rel, oracle_run

pytestmark = pytest.mark.gpu

CASES = [
    # name, kwargs
    ("hcz_cd_periodic_128", dict(model="ac_hcz")),
    ("acp_cd_periodic_128", dict(model="ac_p")),
    ("hcz_cd_walls_64x96_droplet", dict(model="ac_hcz", nx=64, ny=96, walls=True, sign=-1, radius=20.0)),
    ("acp_cd_walls_96x64", dict(model="ac_p", nx=96, ny=64, walls=True, radius=20.0)),
    # non-equilibrium-extrapolation walls, with the interface reaching the wall rows
    ("hcz_cd_nee_64x48_droplet", dict(model="ac_hcz", nx=64, ny=48, walls=True, nee=True, sign=-1, radius=21.0)),
    ("acp_cd_nee_64x48", dict(model="ac_p", nx=64, ny=48, walls=True, nee=True, radius=21.0)),
    ("hcz_uw_periodic_96x64", dict(model="ac_hcz", nx=96, ny=64, flux=abi.MPD_FLUX_UW, radius=20.0)),
    ("acp_uw_periodic_64", dict(model="ac_p", nx=64, ny=64, flux=abi.MPD_FLUX_UW, radius=20.0)),
    ("pp_cd_periodic_64", dict(model="pp_d2q9", nx=64, ny=64)),
    ("pp_uw_periodic_80x48", dict(model="pp_d2q9", nx=80, ny=48, flux=abi.MPD_FLUX_UW, radius=15.0)),
]


def _cmp(gpu, ref, tag):
    planes = [("f", gpu.f, ref.f), ("m", gpu.m, ref.m)]
    if gpu.h is not None:
        planes.append(("h", gpu.h, ref.h))
    if gpu.cfg.model == abi.MPD_MODEL_PP_D2Q9:
        planes += [("ffx", gpu.ffx, ref.ffx), ("ffy", gpu.ffy, ref.ffy)]
    for name, a, b in planes:
        if not np.array_equal(a, b):
            raise AssertionError("%s: plane %s differs, max rel %.3e" % (tag, name, max_rel(a, b)))


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_phase_path_bit_exact_vs_oracle(name, kw, oracle_mod):
    cfg, st = make_case(**kw)
    steps = [1, 5, 25]
    ref = oracle_run(cfg, st, steps)
    s = solver.Solver(cfg, path=abi.MPD_PATH_PHASES)
    s.upload(st)
    done = 0
    for k in steps:
        s.step(k - done)
        done = k
        _cmp(s.download(), ref[k], "%s step %d" % (name, k))
    s.close()


def test_reductions_match_host_sums():
    cfg, st = make_case("ac_hcz", nx=96, ny=80, radius=25.0)
    s = solver.Solver(cfg, path=1)
    s.upload(st)
    s.step(20)
    g = s.download()
    r = s.reductions(update_residual=True)
    rho, u, v = g.m[abi.M["rho"]], g.m[abi.M["u"]], g.m[abi.M["v"]]
    assert r["status"] == 0 and r["nonfinite"] == 0
    assert abs(r["sum_rho"] - rho.sum()) <= 1e-12 * abs(rho.sum())
    assert abs(r["sum_rho_u"] - (rho * u).sum()) <= 1e-12 * np.abs(rho * u).sum() + 1e-300
    assert abs(r["kinetic_energy"] - 0.5 * (rho * (u * u + v * v)).sum()) <= 1e-12 * r["kinetic_energy"] + 1e-300
    # first residual is against Rho_1k = 0, i.e. exactly 1 (DUGKS2DSolver.cpp:815-820)
    assert abs(r["residual"] - 1.0) < 1e-14
    s.close()


def test_nonfinite_state_is_reported():
    cfg, st = make_case("ac_p", nx=32, ny=32, radius=8.0)
    st.h[:, 100] = np.nan
    s = solver.Solver(cfg, path=1)
    s.upload(st)
    s.step(1)
    r = s.reductions()
    assert r["status"] == -3 and r["nonfinite"] >= 1
    s.close()
