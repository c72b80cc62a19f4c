"""The serving loop (mpdugks_amd.serve) with the real GPU engine: Poisson
two-droplet coalescence requests, admitted by an HBM budget small enough to
make requests queue, stepped together in one batch, each retired at a kinetic-
energy residual of 1e-5 (reachable in FP32-bulk, whose residual floor is
1e-6..3e-5) or at the step cap.  A retired request's state equals the same case
stepped alone on the FP32 fused path for the same number of steps, bit for bit.
With an FP64 continuation the requests reach the configuration's 1e-8."""
import numpy as np
import pytest

from mpdugks_amd import abi, cases, serve, solver

pytestmark = pytest.mark.gpu


def test_server_with_gpu_engine():
    reqs = serve.coalescence_requests(4, rate=40.0, seed=5, radii=(8.0, 10.0))
    assert len(reqs) == 4 and all(r.params["side"] == 64 for r in reqs)
    budget = int(2.5 * reqs[0].bytes_estimate)        # two resident at most: the others queue
    eng = serve.GpuEngine(keep_final=True)
    srv = serve.Server(eng, hbm_budget=budget, check_every=1000, ke_tol=1e-5, max_steps=60000)
    stats = srv.run(reqs)
    eng.close()
    assert stats["completed"] == 4 and stats["rejected"] == 0
    assert stats["max_active"] <= 2 and stats["queued_for_budget"] >= 1
    assert stats["converged"] >= 1
    assert stats["p99_latency_s"] >= stats["p50_latency_s"] > 0
    # one retired request against the same case stepped alone
    r = next(q for q in reqs if q.converged)
    cfg, drops = r.make()
    s = solver.Solver(cfg, path=abi.MPD_PATH_FUSED)
    s.init_drops(drops, sign=-1)
    s.step(r.steps)
    alone = s.download()
    s.close()
    assert np.array_equal(alone.f, r.final.f) and np.array_equal(alone.h, r.final.h)
    assert np.array_equal(alone.m, r.final.m)


def test_server_continues_in_fp64_to_1e8():
    """A KE residual of 1e-8 lies below what FP32-bulk resolves (its floor is
    1e-6..3e-5).  Two symmetric coalescences (radii 8.8 and 8.8, gap 4; FP64
    alone settles there, below 1e-12): each reaches 1e-5 in the FP32 batch,
    continues as an FP64 handle from its downloaded state and retires at 1e-8.
    The final state equals the same case stepped alone in FP32 to the switch
    step, downloaded, and stepped in FP64 for the rest, bit for bit."""
    def make():
        return cases.coalescence_config(8.8, 8.8, 4, sigma=1e-2)
    reqs = [serve.Request(k, 0.01 * k, make, bytes_estimate=64 * 64 * 1100, cells=64 * 64) for k in range(2)]
    eng = serve.GpuEngine(keep_final=True)
    srv = serve.Server(eng, hbm_budget=1 << 30, check_every=1000, ke_tol=1e-8, max_steps=150000,
                       switch_tol=1e-5)
    stats = srv.run(reqs)
    eng.close()
    print(stats)
    assert stats["completed"] == 2 and stats["converged"] == 2 and stats["continued_fp64"] == 2
    r = reqs[0]
    cfg, drops = r.make()
    s = solver.Solver(cfg, path=abi.MPD_PATH_FUSED)
    s.init_drops(drops, sign=-1)
    s.step(r.fp64_from)
    mid = s.download()
    s.close()
    cfg64 = type(cfg).from_buffer_copy(cfg)
    cfg64.precision = abi.MPD_FP64
    t = solver.Solver(cfg64, path=abi.MPD_PATH_FUSED)
    t.upload(solver.State(cfg64, mid.h, mid.f, mid.m))
    t.step(r.steps - r.fp64_from)
    alone = t.download()
    t.close()
    assert np.array_equal(alone.f, r.final.f) and np.array_equal(alone.h, r.final.h)
    assert np.array_equal(alone.m, r.final.m)


@pytest.mark.parametrize("precision", [abi.MPD_FP32_BULK, abi.MPD_FP64])
def test_admission_estimate_covers_device_bytes(precision):
    """coalescence_requests budgets 1100 B per cell.  That covers a primed case in
    either precision: the FP64 explicit state, the phase scratch of the priming
    steps and the two derived plane sets (FP32, or FP64 for a continuation)."""
    for side in (64, 256):
        cfg, drops = cases.coalescence_config(10.0, 10.0, 4, precision=precision, sides=(side,))
        s = solver.Solver(cfg, path=abi.MPD_PATH_FUSED)
        s.init_drops(drops, sign=-1)
        s.step(3)
        b = s.device_bytes()
        s.close()
        print("precision %d, %d^2: %.1f B per cell" % (precision, side, b / side ** 2))
        assert b <= 1100 * side * side
