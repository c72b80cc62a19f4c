"""The largest mesh the FP32-bulk Allen-Cahn kernel accepts.

The kernel addresses a state set through 32-bit buffer resources, and
mpd_create admits an HCZ mesh while its 22 float planes span less than 4 GiB:
at most 48,806,446 cells (tests/test_abi_cpu.py refuses the next row up).  On
8192 x 5957 cells (48,799,744: the last plane ends 589,824 bytes below 2^32) a
drop of radius 20 sits at the centre of the 320 x 320 window in the mesh's
last rows and columns, where the loads and stores use the highest offsets.

The drop's tanh profile (width 5) is exactly the vapour value beyond 48 cells
from the interface (69 cells from its centre), every cell beyond that starts
in the same rest state, and a step reads at most 3 cells around each cell.
After 20 steps the drop's influence has reached at most 69 + 60 = 129 cells
from the centre, inside the window's 160, while every cell outside the window
still holds the same uniform state in both meshes; so the window must evolve
exactly as the same drop on a 320 x 320 periodic mesh: bit for bit (the tile
shape, which differs between the two meshes, never changes the state on a
periodic mesh), every plane.
sum(phi) over the whole mesh is conserved to 1e-12.
"""
import gc
import sys

import numpy as np
import pytest

from mpdugks_amd import abi, solver
from mpdugks_amd.config import default_config

pytestmark = pytest.mark.gpu

NX, NY, W, R, STEPS = 8192, 5957, 320, 20.0, 20


def _say(msg):
    print(msg)
    sys.stdout.flush()


def _run(nx, ny, cx, cy):
    cfg = default_config(abi.MPD_MODEL_AC_HCZ, nx=nx, ny=ny)
    cfg.precision = abi.MPD_FP32_BULK
    s = solver.Solver(cfg)
    s.init_drops([(cx, cy, R)], sign=-1)
    m0 = s.reductions()["sum_phi"]
    s.step(STEPS)
    r = s.reductions()
    _say("%dx%d: %d steps, tile %d, sum(phi) %.15e -> %.15e, %d non-finite"
         % (nx, ny, STEPS, s.tile(), m0, r["sum_phi"], r["nonfinite"]))
    g = s.download()
    s.close()
    return g, m0, r


def test_largest_fp32_mesh_matches_small_mesh_window():
    assert NX * NY * 22 * 4 < 2 ** 32 <= NX * (NY + 1) * 22 * 4
    ox, oy = NX - W, NY - W
    h = W // 2
    small, _, _ = _run(W, W, float(h), float(h))
    _say("largest mesh %dx%d, window at (%d, %d)" % (NX, NY, ox, oy))
    big, m0, r = _run(NX, NY, float(ox + h), float(oy + h))
    assert r["status"] == 0 and r["nonfinite"] == 0
    assert abs(r["sum_phi"] / m0 - 1) < 1e-12
    for name in ("h", "f", "m"):
        a = getattr(big, name).reshape(-1, NY, NX)[:, oy:, ox:]
        b = getattr(small, name).reshape(-1, W, W)
        if not np.array_equal(a, b):
            bad = np.argwhere(a != b)
            raise AssertionError("plane set %s differs at %d entries (first %s)" % (name, len(bad), bad[0].tolist()))
    # the window holds the drop, not only the uniform vapour
    phi = small.m[abi.M["phi"]].reshape(W, W)
    assert 0.99 < phi.max() <= 1.0 and phi.min() >= -1e-6
    del big, small
    gc.collect()
