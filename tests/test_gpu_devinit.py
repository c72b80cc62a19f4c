"""Start states built on the GPU against the host initialisers.

The host initialisers are bit-exact with the reference's (test_oracle_golden);
the device ones use the device's tanh and sqrt, so they agree to a few units in
the last place, and a step from either start agrees to the same order.
"""
import numpy as np
import pytest

from mpdugks_amd import abi, cases, solver
from mpdugks_amd.config import default_config

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("model,sign", [("ac_hcz", -1), ("ac_p", +1)])
def test_one_drop_matches_host(model, sign):
    cfg = default_config(abi.MODEL_NAMES[model], nx=96, ny=80)
    host = solver.ac_drop(cfg, cx=40.0, cy=37.5, radius=21.0, sign=sign)
    s = solver.Solver(cfg, path=abi.MPD_PATH_PHASES)
    s.init_drops([(40.0, 37.5, 21.0)], sign=sign)
    dev = s.download()
    for name, a, b in (("h", dev.h, host.h), ("f", dev.f, host.f), ("m", dev.m, host.m)):
        assert _rel(a, b) < 1e-14, name
    # one step from each start
    s.step(3)
    a = s.download()
    t = solver.Solver(cfg, path=abi.MPD_PATH_PHASES)
    t.upload(host)
    t.step(3)
    b = t.download()
    assert _rel(a.f, b.f) < 1e-12 and _rel(a.m[abi.M["rho"]], b.m[abi.M["rho"]]) < 1e-12


def test_two_drops_match_host_coalescence():
    cfg, host = cases.coalescence_case(18.0, 12.0, 4, precision=abi.MPD_FP64)
    drops = cases.coalescence_drops(18.0, 12.0, 4, cfg)
    s = solver.Solver(cfg, path=abi.MPD_PATH_PHASES)
    s.init_drops(drops, sign=-1)
    dev = s.download()
    for name, a, b in (("h", dev.h, host.h), ("f", dev.f, host.f), ("m", dev.m, host.m)):
        assert _rel(a, b) < 1e-14, name


def test_interface_matches_host():
    cfg, host = cases.rayleigh_taylor(nx=64, ny=128, precision=abi.MPD_FP64)
    y0 = cases.rayleigh_taylor_y0(cfg)
    s = solver.Solver(cfg, path=abi.MPD_PATH_PHASES)
    s.init_interface(y0, sign=+1)
    dev = s.download()
    for name, a, b in (("h", dev.h, host.h), ("f", dev.f, host.f), ("m", dev.m, host.m)):
        assert _rel(a, b) < 1e-14, name


@pytest.mark.parametrize("model", ["pp_d2q9", "pp_d2v16"])
def test_pp_noise_is_equilibrium_in_range(model, oracle_mod):
    cfg = default_config(abi.MODEL_NAMES[model], nx=64, ny=48, pp_tr=0.8)
    s = solver.Solver(cfg)
    s.init_pp_noise(0.1645, amp=0.01, seed=5)
    dev = s.download()
    rho = dev.m[abi.M["rho"]]
    assert rho.min() >= 0.1645 * 0.99 and rho.max() < 0.1645 * 1.01
    assert abs(rho.mean() / 0.1645 - 1) < 2e-3 and rho.std() > 0.1645 * 0.004
    # tilde = the equilibrium at rest of each cell's density, bit for bit
    if model == "pp_d2v16":
        for n in (0, 17, 999):
            assert np.array_equal(dev.f[:, n], oracle_mod.d2v16_feq(rho[n], 0.0, 0.0, cfg.rt))
    else:
        w = np.array([4 / 9, 1 / 9, 1 / 36, 1 / 9, 1 / 36, 1 / 9, 1 / 36, 1 / 9, 1 / 36])
        assert np.array_equal(dev.f, w[:, None] * rho[None, :] * 1.0)
    assert np.array_equal(dev.ffx, np.zeros_like(dev.ffx))
    assert np.array_equal(dev.m[abi.M["tau"]], np.full_like(rho, cfg.pp_tau))
    # a different seed gives a different field; the same seed the same one
    s.init_pp_noise(0.1645, amp=0.01, seed=5)
    assert np.array_equal(s.download().m[abi.M["rho"]], rho)
    s.init_pp_noise(0.1645, amp=0.01, seed=6)
    assert not np.array_equal(s.download().m[abi.M["rho"]], rho)


def test_device_start_then_fused_fp32():
    cfg, _ = cases.coalescence_case(15.0, 15.0, 3)
    drops = cases.coalescence_drops(15.0, 15.0, 3, cfg)
    s = solver.Solver(cfg, path=abi.MPD_PATH_FUSED)
    s.init_drops(drops, sign=-1)
    s.step(50)
    r = s.reductions()
    assert r["status"] == 0 and r["nonfinite"] == 0


@pytest.mark.parametrize("model,start", [("pp_d2q9", "drop"), ("pp_d2q9", "flat"), ("pp_d2v16", "drop")])
def test_pp_starts_match_host(model, start):
    """SCMP_Drop and SCMP_FlatInterface built on the device against the host
    initialisers (bit-exact with the reference at step 0), on a handle that
    has stepped before: the face forces and the previous step's face Psi (read
    by the reference face order) are reset as by an upload, so a step from the
    device start equals a step from the uploaded host start to round-off."""
    m = abi.MODEL_NAMES[model]
    cfg = default_config(m, nx=96, ny=80)
    if start == "drop":
        host = solver.scmp_drop(cfg, cx=40.0, cy=37.5, radius=21.0)
    else:
        host = solver.scmp_flat(cfg, y_lo=20.0, y_hi=60.0)
    ref_order = m == abi.MPD_MODEL_PP_D2Q9
    runs = []
    for device in (True, False):
        s = solver.Solver(cfg, path=abi.MPD_PATH_PHASES)
        if ref_order:
            s.set_face_order(abi.MPD_FACE_ORDER_REFERENCE)
        s.init_pp_noise(0.16, amp=0.01, seed=1)
        s.step(3)  # leaves non-zero face forces and face Psi behind
        if not device:
            s.upload(host)
        elif start == "drop":
            s.init_pp_drop(cx=40.0, cy=37.5, radius=21.0)
        else:
            s.init_pp_flat(y_lo=20.0, y_hi=60.0)
        st = s.download()
        if device:
            assert _rel(st.f, host.f) < 1e-14 and _rel(st.m, host.m) < 1e-14
            assert not st.ffx.any() and not st.ffy.any()
            if ref_order:
                assert not s.download_face_psi().any()
        s.step(5)
        runs.append(s.download())
        s.close()
    a, b = runs
    assert _rel(a.f, b.f) < 1e-12 and _rel(a.m[abi.M["rho"]], b.m[abi.M["rho"]]) < 1e-12
    if ref_order:
        assert _rel(a.ffx, b.ffx) < 1e-9
