"""mpd_device_init_interface_noalloc (the allocation-free variant) against
mpd_device_init_interface: the same start state bit for bit, on two calls in a
row (the second on a smaller mesh), the same state after FP64 steps, and a
clear refusal beyond its 16384-column limit."""
import numpy as np
import pytest

from mpdugks_amd import abi, solver
from mpdugks_amd.config import default_config

pytestmark = pytest.mark.gpu


def _y0(nx, seed):
    rng = np.random.default_rng(seed)
    x = np.arange(nx) + 0.5
    return 0.5 * 48 + 4.0 * np.cos(2 * np.pi * x / nx) + rng.uniform(-0.5, 0.5, nx)


def _pair(nx, ny, seed):
    cfg = default_config(abi.MPD_MODEL_AC_HCZ, nx=nx, ny=ny, width=4.0, rho_l=3.0, rho_v=1.0)
    cfg.bc_y = abi.MPD_BC_WALL_BB
    y0 = _y0(nx, seed)
    a = solver.Solver(cfg, path=abi.MPD_PATH_PHASES)
    a.init_interface(y0, sign=+1)
    b = solver.Solver(cfg, path=abi.MPD_PATH_PHASES)
    b.init_interface_noalloc(y0, sign=+1)
    return a, b


def _equal(a, b):
    sa, sb = a.download(), b.download()
    return np.array_equal(sa.h, sb.h) and np.array_equal(sa.f, sb.f) and np.array_equal(sa.m, sb.m)


def test_noalloc_interface_matches_allocating_variant():
    for nx, ny in ((96, 48), (40, 24)):   # the second call on a smaller mesh
        a, b = _pair(nx, ny, seed=nx)
        assert b.device_bytes() == a.device_bytes()
        assert _equal(a, b), (nx, ny)
        a.step(3)
        b.step(3)
        assert _equal(a, b), (nx, ny)
        a.close()
        b.close()


def test_noalloc_interface_refuses_beyond_its_limit():
    cfg = default_config(abi.MPD_MODEL_AC_HCZ, nx=16385, ny=4)
    s = solver.Solver(cfg)
    with pytest.raises(solver.MpdError, match="at most 16384 columns"):
        s.init_interface_noalloc(np.full(16385, 2.0))
    s.close()
