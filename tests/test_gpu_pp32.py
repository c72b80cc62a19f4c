"""FP32-bulk pseudopotential D2Q9 path (pp32.hip), CD and UW faces, against the FP64 oracle.

Populations, bar-plus, face values and fluxes are FP32; the Carnahan-Starling
pressure and Psi are FP64; the global sums are FP64 (compensated).  Stated
tolerances, as for the Allen-Cahn FP32-bulk kernel: density and populations
within 1e-5 relative L2 of the FP64 oracle, velocity within 1e-3 relative to
max(|u|_L2, 1e-6 sqrt(N)).  The density is the conserved variable in FP64
(face mass fluxes, f~_0 the remainder), so the mass sum stays within 1e-9 of
the FP64 oracle's.
"""
import numpy as np
import pytest

from mpdugks_amd import abi, solver
from helpers import make_case, oracle_run
from mpdugks_amd.config import default_config

pytestmark = pytest.mark.gpu


def rel_l2(a, b, floor=0.0):
    return float(np.sqrt(((a - b) ** 2).sum()) / max(np.sqrt((b ** 2).sum()), floor))


@pytest.mark.parametrize("n,steps", [(64, 100), (64, 500), (48, 1500)])
def test_pp32_error_vs_fp64_oracle(n, steps, oracle_mod):
    cfg, st = make_case("pp_d2q9", nx=n, ny=n, precision=abi.MPD_FP32_BULK)
    ref = oracle_run(cfg, st, [steps])[steps]
    s = solver.Solver(cfg)
    s.upload(st)
    s.step(steps)
    g = s.download()
    r = s.reductions()
    s.close()
    N = n * n
    errs = {"rho": rel_l2(g.m[abi.M["rho"]], ref.m[abi.M["rho"]]), "f": rel_l2(g.f, ref.f),
            "psi": rel_l2(g.m[abi.M["psi"]], ref.m[abi.M["psi"]])}
    for name in ("u", "v"):
        errs[name] = rel_l2(g.m[abi.M[name]], ref.m[abi.M[name]], floor=1e-6 * np.sqrt(N))
    mass = abs(r["sum_rho"] - ref.m[abi.M["rho"]].sum()) / ref.m[abi.M["rho"]].sum()
    print("pp32 %dx%d %d steps: %s mass %.2e" % (n, n, steps, errs, mass))
    assert r["status"] == 0 and r["nonfinite"] == 0
    assert errs["rho"] < 1e-5 and errs["f"] < 1e-5 and errs["psi"] < 1e-5, errs
    assert errs["u"] < 1e-3 and errs["v"] < 1e-3, errs
    assert mass < 1e-9


def test_pp32_state_carries_across_calls():
    """The FP32 planes are the state: stepping 300 + 200 with a download and a
    reduction in between gives the same bits as 500 in one call."""
    cfg, st = make_case("pp_d2q9", nx=40, ny=56, radius=12.0, precision=abi.MPD_FP32_BULK)
    a = solver.Solver(cfg)
    a.upload(st)
    a.step(500)
    ga = a.download()
    a.close()
    b = solver.Solver(cfg)
    b.upload(st)
    b.step(300)
    b.download()
    b.reductions()
    b.step(200)
    gb = b.download()
    b.close()
    for name in ("f", "m", "ffx", "ffy"):
        assert np.array_equal(getattr(ga, name), getattr(gb, name)), name


def test_pp32_device_noise_start_runs():
    """Noisy liquid-vapour start built on the device (T/Tc = 0.85, the
    reference's shipped temperature): the FP32 path keeps the mass and stays
    finite while the phases separate (the density spread grows from 0.02 rho0
    to more than 0.2 rho0 in 2000 steps)."""
    from mpdugks_amd.config import default_config
    cfg = default_config(abi.MPD_MODEL_PP_D2Q9, nx=64, ny=64, pp_tr=0.85)
    cfg.precision = abi.MPD_FP32_BULK
    s = solver.Solver(cfg)
    s.init_pp_noise(0.1645, amp=0.01, seed=11)
    m0 = s.reductions()["sum_rho"]
    s.step(2000)
    r = s.reductions()
    rho = s.download().m[abi.M["rho"]]
    s.close()
    assert r["status"] == 0 and r["nonfinite"] == 0
    assert abs(r["sum_rho"] - m0) / m0 < 1e-9
    assert rho.max() - rho.min() > 10 * 0.02 * 0.1645  # grown from the +-1 % noise


@pytest.mark.parametrize("n", [64, 40])
def test_pp32_one_launch_equals_per_phase(n):
    """The one-launch tile kernel against the per-phase FP32 kernels (same
    device functions, same inputs): equal to FP32 rounding after 300 steps,
    with the phase path selectable on the same FP32 state."""
    cfg, st = make_case("pp_d2q9", nx=n, ny=n + 8, radius=0.25 * n, precision=abi.MPD_FP32_BULK)
    out = {}
    for path in (abi.MPD_PATH_FUSED, abi.MPD_PATH_PHASES):
        s = solver.Solver(cfg, path=path)
        s.upload(st)
        s.step(300)
        out[path] = s.download()
        s.close()
    a, b = out[abi.MPD_PATH_FUSED], out[abi.MPD_PATH_PHASES]
    for name in ("f", "ffx", "ffy"):
        x, y = getattr(a, name), getattr(b, name)
        err = float(np.abs(x - y).max() / np.abs(y).max())
        print("pp32 one-launch vs per-phase %dx%d %s: max rel %.3e" % (n, n + 8, name, err))
        assert err < 1e-5, (name, err)
    m = rel_l2(a.m[abi.M["rho"]], b.m[abi.M["rho"]])
    assert m < 1e-6, m


def test_pp32_switch_paths_mid_run():
    cfg, st = make_case("pp_d2q9", nx=48, ny=48, precision=abi.MPD_FP32_BULK)
    ref = solver.Solver(cfg, path=abi.MPD_PATH_FUSED)
    ref.upload(st)
    ref.step(200)
    r = ref.download()
    ref.close()
    s = solver.Solver(cfg, path=abi.MPD_PATH_FUSED)
    s.upload(st)
    s.step(70)
    s.set_path(abi.MPD_PATH_PHASES)
    s.step(60)
    s.set_path(abi.MPD_PATH_FUSED)
    s.step(70)
    g = s.download()
    s.close()
    assert rel_l2(g.f, r.f) < 1e-6


def test_pp32_tile_shapes_agree_bit_for_bit():
    """The one-launch step with 16x15 and with 32x16 tiles (a launch shape
    only): the same state bit for bit, on a mesh neither shape fills."""
    out = []
    for tile in (abi.MPD_TILE_16X15, abi.MPD_TILE_32X16):
        cfg = default_config(abi.MPD_MODEL_PP_D2Q9, nx=100, ny=70, pp_tr=0.85)
        cfg.precision = abi.MPD_FP32_BULK
        s = solver.Solver(cfg)
        s.set_tile(tile)
        assert s.tile() == tile
        s.init_pp_noise(0.1645, amp=0.01, seed=5)
        s.step(120)
        out.append(s.download())
        s.close()
    a, b = out
    assert np.array_equal(a.f, b.f) and np.array_equal(a.m, b.m) and np.array_equal(a.ffx, b.ffx)
    cfg = default_config(abi.MPD_MODEL_PP_D2Q9, nx=512, ny=512, pp_tr=0.85)
    cfg.precision = abi.MPD_FP32_BULK
    s = solver.Solver(cfg)
    assert s.tile() == abi.MPD_TILE_32X16      # 512 tiles of 32x16: every CU gets two
    with pytest.raises(solver.MpdError):
        s.set_tile(abi.MPD_TILE_16X16)
    s.close()


@pytest.mark.parametrize("nx,ny,steps", [(64, 64, 100), (64, 64, 500), (70, 47, 300), (20, 18, 200)],
                         ids=["64_100", "64_500", "70x47_300", "20x18_200"])
def test_pp32_uw_error_vs_fp64_oracle(nx, ny, steps, oracle_mod):
    """UW faces on the one-launch tile kernel (16x15 tiles; the bar-plus and Psi
    regions one cell wider for the 6-point gradients) against the FP64
    oracle's UW step, at the bar of the CD cases; tile-multiple, ragged and
    narrower-than-a-tile meshes.  The result is far from the oracle's CD run
    (the face scheme takes effect)."""
    cfg, st = make_case("pp_d2q9", nx=nx, ny=ny, radius=0.25 * min(nx, ny), flux=abi.MPD_FLUX_UW,
                        precision=abi.MPD_FP32_BULK)
    ref = oracle_run(cfg, st, [steps])[steps]
    cd = cfg.__class__.from_buffer_copy(cfg)
    cd.flux = abi.MPD_FLUX_CD
    ref_cd = oracle_run(cd, st, [steps])[steps]
    s = solver.Solver(cfg)
    assert s.tile() == abi.MPD_TILE_16X15
    with pytest.raises(solver.MpdError):
        s.set_path(abi.MPD_PATH_PHASES)
    s.upload(st)
    s.step(steps)
    g = s.download()
    r = s.reductions()
    s.close()
    N = nx * ny
    errs = {"rho": rel_l2(g.m[abi.M["rho"]], ref.m[abi.M["rho"]]), "f": rel_l2(g.f, ref.f),
            "psi": rel_l2(g.m[abi.M["psi"]], ref.m[abi.M["psi"]])}
    for name in ("u", "v"):
        errs[name] = rel_l2(g.m[abi.M[name]], ref.m[abi.M[name]], floor=1e-6 * np.sqrt(N))
    d_cd = rel_l2(g.f, ref_cd.f)
    mass = abs(r["sum_rho"] - ref.m[abi.M["rho"]].sum()) / ref.m[abi.M["rho"]].sum()
    print("pp32 UW %dx%d %d steps: %s mass %.2e; f vs the CD oracle %.2e" % (nx, ny, steps, errs, mass, d_cd))
    assert r["status"] == 0 and r["nonfinite"] == 0
    assert errs["rho"] < 1e-5 and errs["f"] < 1e-5 and errs["psi"] < 1e-5, errs
    assert errs["u"] < 1e-3 and errs["v"] < 1e-3, errs
    assert mass < 1e-9
    assert d_cd > 100 * errs["f"]
