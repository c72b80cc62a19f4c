"""FP32-bulk with UW faces (UW_Interior_phi_Bh, Ark/Cpp/DUGKS2DSolver.cpp:560-603,
with the 6-point gradients of Grad_VS_6points, GradSchemeBasic.h:95-126) against
the FP64 oracle's UW step (bit-exact with the reference after one step,
tests/test_oracle_golden.py): the FP32 bar of the CD cases -- rho, phi, p
within 1e-5 relative L2, velocity within 1e-3 of max(|u|_L2, 1e-6 sqrt(N)) --
on horizons where the FP64 problem is well conditioned; tile-multiple,
ragged and narrower-than-tile-plus-halo meshes, periodic and with bounce-back
and NEE walls.  The UW result is far
closer to the oracle's UW run than to its CD run (the face scheme takes
effect), and sum(phi) is conserved."""
import numpy as np
import pytest

from mpdugks_amd import abi, solver
from helpers import make_case, oracle_run

pytestmark = pytest.mark.gpu


def rel_l2(a, b, floor=0.0):
    return float(np.sqrt(((a - b) ** 2).sum()) / max(np.sqrt((b ** 2).sum()), floor))


# Meshes narrower than tile + halo (nx < 22 or ny < 21 for the 16x15 tile with
# its 3-cell UW halo) take the kernel's general periodic wrap.  They are chosen
# well conditioned: the FP64 run from a start perturbed by 6e-8 relative moves
# rho, phi and p by at most 3e-9 over the horizon (a drop of radius 4 on 19x13,
# narrower than its interface, is not: there the perturbed FP64 run moves phi
# by 5e-4 within 10 steps, CD and UW alike).
CASES = [
    # model, nx, ny, radius, sign, y boundary, steps
    ("ac_hcz", 96, 96, 30.0, 1, "periodic", 60),
    ("ac_p", 96, 96, 30.0, 1, "periodic", 500),
    ("ac_hcz", 70, 47, 14.0, 1, "periodic", 60),     # not a multiple of the 16x15 tile
    ("ac_p", 20, 16, 6.0, -1, "periodic", 100),      # narrower than tile + halo both ways
    ("ac_hcz", 21, 40, 8.0, 1, "periodic", 100),     # narrower in x only
    ("ac_p", 18, 44, 7.0, 1, "periodic", 200),
    # walls: ghost rows in the 6-point gradients, the shadow side of a wall face
    # with gradient 0 (perturbed-FP64 distances <= 3e-9 over these horizons)
    ("ac_hcz", 64, 48, 21.0, -1, "bb", 60),          # droplet reaching 3 cells from both walls
    ("ac_p", 70, 47, 20.0, 1, "nee", 300),
    ("ac_hcz", 70, 47, 15.0, 1, "nee", 60),
    ("ac_p", 20, 16, 5.0, -1, "bb", 100),            # one tile row holding both walls
    ("ac_hcz", 24, 30, 8.0, -1, "nee", 100),
]


@pytest.mark.parametrize("model,nx,ny,radius,sign,bc,steps", CASES,
                         ids=["%s_%dx%d_%s" % (c[0], c[1], c[2], c[5]) for c in CASES])
def test_fp32_uw_vs_fp64_oracle(model, nx, ny, radius, sign, bc, steps, oracle_mod):
    cfg, st = make_case(model, nx=nx, ny=ny, radius=radius, sign=sign, flux=abi.MPD_FLUX_UW,
                        precision=abi.MPD_FP32_BULK, walls=bc != "periodic", nee=bc == "nee")
    ref = oracle_run(cfg, st, [steps])[steps]
    cd = cfg.__class__.from_buffer_copy(cfg)
    cd.flux = abi.MPD_FLUX_CD
    ref_cd = oracle_run(cd, st, [steps])[steps]
    s = solver.Solver(cfg)
    s.upload(st)
    s.step(steps)
    assert s.tile() == abi.MPD_TILE_16X15
    g = s.download()
    s.close()
    N = nx * ny
    errs = {k: rel_l2(g.m[abi.M[k]], ref.m[abi.M[k]]) for k in ("rho", "phi", "p")}
    errs.update({k: rel_l2(g.m[abi.M[k]], ref.m[abi.M[k]], floor=1e-6 * np.sqrt(N)) for k in ("u", "v")})
    d_cd = rel_l2(g.m[abi.M["phi"]], ref_cd.m[abi.M["phi"]])
    print("fp32 UW %s %dx%d %s %d steps: %s; phi vs the CD oracle %.3e" % (model, nx, ny, bc, steps, errs, d_cd))
    assert errs["rho"] < 1e-5 and errs["phi"] < 1e-5 and errs["p"] < 1e-5, errs
    assert errs["u"] < 1e-3 and errs["v"] < 1e-3, errs
    assert d_cd > 100 * errs["phi"]
    assert abs(g.m[abi.M["phi"]].sum() / st.m[abi.M["phi"]].sum() - 1) < (1e-12 if bc == "periodic" else 1e-11)


def test_fp32_uw_refuses_other_tiles():
    cfg, _ = make_case("ac_hcz", nx=32, ny=32, flux=abi.MPD_FLUX_UW, precision=abi.MPD_FP32_BULK)
    s = solver.Solver(cfg)
    for t in (abi.MPD_TILE_16X16, abi.MPD_TILE_32X16):
        with pytest.raises(solver.MpdError):
            s.set_tile(t)
    s.close()
