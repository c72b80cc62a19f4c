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
    # UW with walls: ghost rows in the 6-point gradients, the wall face's shadow side
    ("hcz_uw_walls_64x48_droplet", dict(model="ac_hcz", nx=64, ny=48, flux=abi.MPD_FLUX_UW, walls=True, sign=-1,
                                        radius=21.0)),
    ("hcz_uw_nee_70x47", dict(model="ac_hcz", nx=70, ny=47, flux=abi.MPD_FLUX_UW, walls=True, nee=True,
                              radius=20.0)),
    ("acp_uw_nee_64x48", dict(model="ac_p", nx=64, ny=48, flux=abi.MPD_FLUX_UW, walls=True, nee=True,
                              radius=21.0)),
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


@pytest.mark.parametrize("name", ["pp", "pp_shift", "pp_flat", "pp_uw"])
def test_pp_reference_face_order_bit_exact_vs_reference(name):
    """The FP64 pseudopotential step in the reference's one-thread face order
    reproduces the compiled reference's own state digests (tests/golden,
    OMP_NUM_THREADS=1) after 1, 2, 10 and 100 steps, bit for bit."""
    from test_oracle_golden import check_state, load, pp_case
    g = load(name)
    cfg, st = pp_case(g)
    s = solver.Solver(cfg, path=abi.MPD_PATH_PHASES)
    s.set_face_order(abi.MPD_FACE_ORDER_REFERENCE)
    s.upload(st)
    done = 0
    for k in sorted(int(x) for x in g["steps"] if int(x) > 0):
        s.step(k - done)
        done = k
        bad = check_state(s.download(), g["steps"][str(k)])
        assert bad == [], "%s step %d: planes %s" % (name, k, bad)
    s.close()


@pytest.mark.parametrize("flux", [abi.MPD_FLUX_CD, abi.MPD_FLUX_UW], ids=["cd", "uw"])
def test_pp_reference_face_order_vs_oracle_nonsquare(flux, oracle_mod):
    """Reference face order on a 48x40 mesh with the drop across a corner: the
    face ranks depend on nx and ny separately; GPU state, face forces and face
    Psi equal the oracle's reference-order mode bit for bit (with UW faces the
    second copies of the periodic faces, which the drop crosses, have their
    own Psi and force)."""
    cfg, st = make_case("pp_d2q9", nx=48, ny=40, radius=12.0, flux=flux)
    sx, sy = 20, 17
    roll = lambda a: np.roll(a.reshape(a.shape[0], 40, 48), (sy, sx), axis=(1, 2)).reshape(a.shape)
    st = solver.State(cfg, None, roll(st.f), roll(st.m), st.ffx, st.ffy)
    o = oracle_mod.Oracle(cfg)
    o.set_state(np.zeros_like(st.f), st.f, st.m, st.ffx, st.ffy)
    o.set_face_order(True)
    s = solver.Solver(cfg, path=abi.MPD_PATH_PHASES)
    s.set_face_order(abi.MPD_FACE_ORDER_REFERENCE)
    s.upload(st)
    done = 0
    for k in (1, 3, 20):
        o.step(k - done)
        s.step(k - done)
        done = k
        h, f, m, ffx, ffy = o.get_state()
        _cmp(s.download(), solver.State(cfg, None, f, m, ffx, ffy), "48x40 step %d" % k)
        assert np.array_equal(s.download_face_psi(), o.get_face_extra()[0])
    s.close()


def test_pp_face_psi_resume_matches_continuous_run():
    """Stopping after 5 steps, downloading the state and the face Psi and
    resuming in a new handle gives the continuous 10-step run bit for bit."""
    cfg, st = make_case("pp_d2q9", nx=64, ny=64)
    a = solver.Solver(cfg, path=abi.MPD_PATH_PHASES)
    a.set_face_order(abi.MPD_FACE_ORDER_REFERENCE)
    a.upload(st)
    a.step(10)
    ref = a.download()
    b = solver.Solver(cfg, path=abi.MPD_PATH_PHASES)
    b.set_face_order(abi.MPD_FACE_ORDER_REFERENCE)
    b.upload(st)
    b.step(5)
    mid, psi = b.download(), b.download_face_psi()
    c = solver.Solver(cfg, path=abi.MPD_PATH_PHASES)
    c.set_face_order(abi.MPD_FACE_ORDER_REFERENCE)
    c.upload(mid)
    c.upload_face_psi(psi)
    c.step(5)
    _cmp(c.download(), ref, "resumed")
    for x in (a, b, c):
        x.close()


@pytest.mark.parametrize("name", ["hcz_uw", "acp_uw", "hcz_uw_wall", "acp_uw_wall", "hcz_uw_nee", "acp_uw_nee"])
def test_uw_bit_exact_vs_reference(name):
    """The FP64 UW step (periodic, bounce-back and NEE walls) reproduces the
    compiled reference's own state digests after 1, 2, 10 and 100 steps, bit
    for bit: the second stored copy of each periodic face included."""
    from test_oracle_golden import check_state, golden_case, load
    g = load(name)
    cfg, st = golden_case(g)
    s = solver.Solver(cfg, path=abi.MPD_PATH_PHASES)
    s.upload(st)
    done = 0
    for k in sorted(int(x) for x in g["steps"] if int(x) > 0):
        s.step(k - done)
        done = k
        bad = check_state(s.download(), g["steps"][str(k)])
        assert bad == [], "%s step %d: planes %s" % (name, k, bad)
    s.close()
