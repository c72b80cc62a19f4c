"""Batched ensemble (one launch per step over many uneven cases) against the
same cases stepped one at a time on the fused FP32 path: the batched kernel
runs the same tile code, so the states must agree bit for bit."""
import numpy as np
import pytest

from mpdugks_amd import abi, cases, solver
from mpdugks_amd.config import default_config

pytestmark = pytest.mark.gpu


def _primed(spec):
    cfg, st = cases.ensemble_case(spec)
    s = solver.Solver(cfg, path=abi.MPD_PATH_FUSED)
    s.upload(st)
    s.step(2)
    return s


def test_batch_matches_individual_cases():
    specs = cases.ensemble_specs(6, seed=3, sides=(32, 48, 64))
    single = [_primed(sp) for sp in specs]
    batched = [_primed(sp) for sp in specs]
    b = solver.Batch()
    for s in batched:
        b.add(s)
    assert b.size()[0] == 6
    b.step(7)          # odd: graph pairs + one direct launch
    b.step(10)
    for s in single:
        s.step(17)
    for a, c in zip(batched, single):
        ga, gc = a.download(), c.download()
        assert np.array_equal(ga.f, gc.f) and np.array_equal(ga.h, gc.h) and np.array_equal(ga.m, gc.m)
    # a member leaves, the rest continue
    b.remove(batched[0])
    b.step(4)
    for s in single[1:]:
        s.step(4)
    for a, c in zip(batched[1:], single[1:]):
        assert np.array_equal(a.download().f, c.download().f)
    b.close()


def test_member_cannot_step_alone():
    s = _primed(cases.ensemble_specs(1, seed=1, sides=(32,))[0])
    b = solver.Batch()
    b.add(s)
    with pytest.raises(solver.MpdError):
        s.step(1)
    # nor be uploaded to or switched to another path while the batch owns it
    st = s.download()
    with pytest.raises(solver.MpdError, match="MPD_ERR_STATE"):
        s.upload(st)
    with pytest.raises(solver.MpdError, match="MPD_ERR_STATE"):
        s.set_path(abi.MPD_PATH_PHASES)
    # nor given a tile shape of its own (the batch has one)
    with pytest.raises(solver.MpdError, match="MPD_ERR_STATE"):
        s.set_tile(abi.MPD_TILE_16X15)
    b.remove(s)
    s.step(1)
    s.upload(st)
    s.set_tile(abi.MPD_TILE_16X15)
    b.close()


def test_unprimed_case_is_rejected():
    cfg, st = cases.ensemble_case(cases.ensemble_specs(1, seed=2, sides=(32,))[0])
    s = solver.Solver(cfg)
    s.upload(st)
    b = solver.Batch()
    with pytest.raises(solver.MpdError):
        b.add(s)
    b.close()


def test_batch_nee_walls_matches_individual_cases():
    """Cases with NEE walls in y batch like periodic ones (same tile code)."""
    def primed(side, r):
        cfg = default_config(abi.MPD_MODEL_AC_HCZ, nx=side, ny=side + 5)
        cfg.precision = abi.MPD_FP32_BULK
        cfg.bc_y = abi.MPD_BC_WALL_NEE
        st = solver.ac_drop(cfg, radius=r, sign=-1)
        s = solver.Solver(cfg, path=abi.MPD_PATH_FUSED)
        s.upload(st)
        s.step(2)
        return s
    dims = [(32, 9.0), (48, 15.0), (40, 12.0)]
    single = [primed(*d) for d in dims]
    batched = [primed(*d) for d in dims]
    b = solver.Batch()
    for s in batched:
        b.add(s)
    b.step(9)
    for s in single:
        s.step(9)
    for a, c in zip(batched, single):
        ga, gc = a.download(), c.download()
        assert np.array_equal(ga.f, gc.f) and np.array_equal(ga.h, gc.h) and np.array_equal(ga.m, gc.m)
    b.close()


def test_batch_tile_shapes_and_autotuner():
    """Forced 16x15, 16x16 and 32x16 batches give the same states bit for bit;
    the autotuner's 24 tuning steps are ordinary steps (the tuned batch equals
    the cases stepped alone) and it keeps the fastest shape."""
    specs = cases.ensemble_specs(5, seed=11, sides=(64, 96, 128))
    runs = {}
    for mode in (abi.MPD_TILE_16X15, abi.MPD_TILE_16X16, abi.MPD_TILE_32X16, abi.MPD_TILE_AUTO):
        members = [_primed(sp) for sp in specs]
        b = solver.Batch()
        for s in members:
            b.add(s)
        b.set_tile(mode)
        b.step(40)
        b.step(3)
        tile, ms = b.tile()
        if mode:
            assert tile == mode
        else:
            assert min(ms) > 0 and tile == 1 + int(np.argmin(ms)), (tile, ms)
        runs[mode] = [m.download() for m in members]
        b.close()
    single = [_primed(sp) for sp in specs]
    for s in single:
        s.step(43)
    ref = [s.download() for s in single]
    for mode, got in runs.items():
        for g, r in zip(got, ref):
            assert np.array_equal(g.f, r.f) and np.array_equal(g.h, r.h) and np.array_equal(g.m, r.m), mode


def test_batch_nee_keeps_16x15_unless_forced():
    cfg = default_config(abi.MPD_MODEL_AC_P, nx=48, ny=40)
    cfg.precision = abi.MPD_FP32_BULK
    cfg.bc_y = abi.MPD_BC_WALL_NEE
    s = solver.Solver(cfg, path=abi.MPD_PATH_FUSED)
    s.upload(solver.ac_drop(cfg, radius=10.0))
    s.step(2)
    b = solver.Batch()
    b.add(s)
    b.step(40)
    assert b.tile() == (abi.MPD_TILE_16X15, [0.0, 0.0, 0.0])
    b.set_tile(abi.MPD_TILE_32X16)
    b.step(2)
    assert b.tile()[0] == abi.MPD_TILE_32X16
    with pytest.raises(solver.MpdError):
        b.set_tile(4)
    b.close()
