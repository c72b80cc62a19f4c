"""Tile shapes of the one-launch Allen-Cahn step.

FP32-bulk step with 16x15, 16x16 and 32x16 tiles: the tile shape
is a launch shape only (every cell and face is computed by the same source
code whichever tile holds it), on meshes that fill no tile shape, with
periodic, BB and NEE y boundaries.  Periodic and BB: the same state bit for
bit.  NEE: the kernel instantiations contract the wall-row arithmetic into
FMAs differently, so the states agree to FP32 rounding (relative L2 within
1e-6 after 30 steps); each is held to the FP64 oracle in
test_gpu_fused_path.py.

FP64 step: 16x15 tiles only; other shapes are refused."""
import numpy as np
import pytest

from mpdugks_amd import abi, cases, solver
from helpers import make_case
from mpdugks_amd.config import default_config

pytestmark = pytest.mark.gpu

CASES = [
    ("rt_256x1024_bb", None),
    ("hcz_periodic_70x45", dict(model="ac_hcz", nx=70, ny=45, radius=15.0)),
    ("hcz_bb_64x48", dict(model="ac_hcz", nx=64, ny=48, walls=True, sign=-1, radius=14.0)),
    ("acp_nee_70x47", dict(model="ac_p", nx=70, ny=47, walls=True, nee=True, radius=14.0)),
    ("hcz_nee_40x18", dict(model="ac_hcz", nx=40, ny=18, walls=True, nee=True, sign=-1, radius=6.0)),
    ("hcz_periodic_8x6", dict(model="ac_hcz", nx=8, ny=6, radius=2.5)),
]


def _run(name, kw, tile, steps):
    if kw is None:
        cfg, st = cases.rayleigh_taylor(nx=256, ny=1024, precision=abi.MPD_FP32_BULK)
    else:
        cfg, st = make_case(precision=abi.MPD_FP32_BULK, **kw)
    s = solver.Solver(cfg, path=abi.MPD_PATH_FUSED)
    s.set_tile(tile)
    assert s.tile() == tile
    s.upload(st)
    s.step(steps)
    g = s.download()
    s.close()
    return g


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_tile_shapes_agree_bit_for_bit(name, kw):
    steps = 30
    a = _run(name, kw, abi.MPD_TILE_16X15, steps)
    for b in (_run(name, kw, abi.MPD_TILE_16X16, steps), _run(name, kw, abi.MPD_TILE_32X16, steps)):
        _agree(name, kw, a, b)


def _agree(name, kw, a, b):
    nee = kw is not None and kw.get("nee", False)
    for plane in ("h", "f", "m"):
        x, y = getattr(a, plane), getattr(b, plane)
        if x is None:
            continue
        if nee:
            d = float(np.sqrt(((x - y) ** 2).sum()) / np.sqrt((y ** 2).sum()))
            print("%s plane %s: relative L2 %.3e between tile shapes" % (name, plane, d))
            assert d < 1e-6, (name, plane, d)
        elif not np.array_equal(x, y):
            bad = np.argwhere(x != y)
            raise AssertionError("%s: plane %s differs at %d entries (first %s)" % (name, plane, len(bad), bad[0].tolist()))


def test_tile_choice_and_validation():
    cfg, st = cases.rayleigh_taylor(nx=256, ny=1024, precision=abi.MPD_FP32_BULK)
    s = solver.Solver(cfg)
    # 256x1024 on a 256-CU device: 1,104 tiles of 16x15 exceed one round of
    # resident workgroups (1,024), the 512 tiles of 32x16 fit one (2 per CU)
    assert s.tile() == abi.MPD_TILE_32X16
    with pytest.raises(solver.MpdError):
        s.set_tile(4)
    s.set_tile(abi.MPD_TILE_16X16)
    assert s.tile() == abi.MPD_TILE_16X16
    s.set_tile(abi.MPD_TILE_AUTO)
    assert s.tile() == abi.MPD_TILE_32X16
    s.close()
    for (nx, ny), want in (((128, 128), abi.MPD_TILE_16X15),      # one round with any shape: the smallest
                           ((1024, 1024), abi.MPD_TILE_16X15),    # several rounds, state within the cache
                           ((1280, 1280), abi.MPD_TILE_32X16)):   # several rounds, state beyond 256 MiB
        cfg = default_config(abi.MPD_MODEL_AC_HCZ, nx=nx, ny=ny)
        cfg.precision = abi.MPD_FP32_BULK
        s = solver.Solver(cfg)
        assert s.tile() == want, (nx, ny, s.tile())
        s.close()


def test_fp64_tile_validation():
    """The FP64 step has 16x15 and 16x16 tiles (two workgroups per CU by LDS).
    256x1024 on a 256-CU device: 1,104 tiles of 16x15 take three rounds of the
    512 slots, 1,024 tiles of 16x16 two, so the host picks 16x16; where both
    take the same rounds it keeps 16x15."""
    cfg, st = cases.rayleigh_taylor(nx=256, ny=1024, precision=abi.MPD_FP64)
    s = solver.Solver(cfg, path=abi.MPD_PATH_FUSED)
    assert s.tile() == abi.MPD_TILE_16X16
    with pytest.raises(solver.MpdError):
        s.set_tile(abi.MPD_TILE_32X16)
    s.set_tile(abi.MPD_TILE_16X15)
    assert s.tile() == abi.MPD_TILE_16X15
    s.set_tile(abi.MPD_TILE_AUTO)
    assert s.tile() == abi.MPD_TILE_16X16
    s.close()
    for (nx, ny), want in (((128, 128), abi.MPD_TILE_16X15),     # one round either way
                           ((768, 768), abi.MPD_TILE_16X15),     # 2,496 / 2,304 tiles: five rounds either way
                           ((1024, 1024), abi.MPD_TILE_16X16)):  # 4,416 / 4,096 tiles: nine / eight rounds
        cfg = default_config(abi.MPD_MODEL_AC_HCZ, nx=nx, ny=ny)
        s = solver.Solver(cfg, path=abi.MPD_PATH_FUSED)
        assert s.tile() == want, (nx, ny, s.tile())
        s.close()
