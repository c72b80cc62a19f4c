"""Kernel timing of a handle (mpd_set_kernel_timing / mpd_kernel_timing).

Region timing on the fused path counts the launches between the end of each
call's first step kernel and the end of its last; priming steps after an
upload and a call's first kernel are not counted, a one-step call is timed
whole.  Per-launch timing counts every step kernel.  The state does not
depend on the timing mode.
"""
import numpy as np
import pytest

from mpdugks_amd import abi, solver
from helpers import make_case

pytestmark = pytest.mark.gpu


def _solver(precision):
    cfg, st = make_case("ac_hcz", nx=64, ny=48, walls=True, sign=-1, radius=12.0, precision=precision)
    s = solver.Solver(cfg, path=abi.MPD_PATH_FUSED)
    s.upload(st)
    return s


@pytest.mark.parametrize("precision", [abi.MPD_FP32_BULK, abi.MPD_FP64])
def test_region_timing_counts(precision):
    s = _solver(precision)
    try:
        s.set_kernel_timing(False)
        s.step(2)          # the two priming steps after an upload
        s.synchronize()
        assert s.kernel_timing()[1] == 0
        s.step(7)
        s.step(1)
        s.synchronize()
        ms, n = s.kernel_timing()
        assert n == 6 + 1
        assert 0.0 < ms < 50.0
        s.step(3)
        ms, n = s.kernel_timing()   # resolves the pending pair itself
        assert n == 2 and ms > 0.0
        assert s.kernel_timing() == (0.0, 0)
    finally:
        s.close()


def test_timing_modes_same_state():
    a, b = _solver(abi.MPD_FP32_BULK), _solver(abi.MPD_FP32_BULK)
    try:
        a.set_kernel_timing(False)
        b.set_kernel_timing(True)
        for s in (a, b):
            s.step(3)
            s.step(1)
            s.step(6)
            s.synchronize()
        assert b.kernel_timing()[1] == 8   # every step kernel after the two priming steps
        ra, rb = a.download(), b.download()
        for name in ("h", "f", "m"):
            assert np.array_equal(getattr(ra, name), getattr(rb, name)), name
    finally:
        a.close()
        b.close()
