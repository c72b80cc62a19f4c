"""Serving-loop logic on the CPU with a stand-in engine (no GPU): admission by
HBM budget, rejection of cases that can never fit, queueing, the KE-residual
stop and the step cap, latency bookkeeping."""
import numpy as np

from mpdugks_amd.serve import Request, Server


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(dt, 1e-4)


class FakeEngine:
    """KE of case r decays as exp(-rate*steps) towards a floor; each step of
    the batch costs 1e-6 s per resident cell on the fake clock."""

    def __init__(self, clock):
        self.clock = clock
        self.resident = {}
        self.max_resident_bytes = 0
        self.used = 0

    def admit(self, r):
        self.resident[r.rid] = r
        self.used += r.bytes_estimate
        self.max_resident_bytes = max(self.max_resident_bytes, self.used)
        r.steps = 2
        return r.bytes_estimate

    def step(self, n):
        self.clock.t += 1e-9 * n * sum(r.cells for r in self.resident.values())

    def kinetic_energy(self, r):
        return 1e-8 + np.exp(-r.cells * 1e-7 * r.steps)

    def retire(self, r):
        del self.resident[r.rid]
        self.used -= r.bytes_estimate


def make_requests(n, gap=0.01, cells=1000, nbytes=100):
    return [Request(k, k * gap, None, nbytes, cells) for k in range(n)]


def test_budget_limits_resident_cases_and_all_complete():
    clk = FakeClock()
    eng = FakeEngine(clk)
    srv = Server(eng, hbm_budget=250, check_every=1000, ke_tol=1e-8, max_steps=10 ** 6,
                 clock=clk, sleep=clk.sleep)
    stats = srv.run(make_requests(10))
    assert stats["completed"] == 10 and stats["rejected"] == 0
    assert eng.max_resident_bytes <= 250          # never more than two 100-byte cases
    assert stats["p99_latency_s"] >= stats["p50_latency_s"] > 0


def test_case_larger_than_budget_is_rejected():
    clk = FakeClock()
    eng = FakeEngine(clk)
    reqs = make_requests(3)
    reqs[1].bytes_estimate = 10 ** 9
    stats = Server(eng, hbm_budget=1000, clock=clk, sleep=clk.sleep).run(reqs)
    assert stats["rejected"] == 1 and stats["completed"] == 2


def test_step_cap_retires_unconverged_cases():
    clk = FakeClock()
    eng = FakeEngine(clk)
    eng.kinetic_energy = lambda r: 1.0 + 0.5 * np.sin(r.steps)   # never settles
    stats = Server(eng, hbm_budget=10 ** 6, check_every=1000, max_steps=5000,
                   clock=clk, sleep=clk.sleep).run(make_requests(4))
    assert stats["completed"] == 4 and stats["converged"] == 0


class FloorEngine(FakeEngine):
    """KE relaxes as 1 + exp(-steps/2000); in FP32 it carries a relative
    rounding jitter of 5e-6 (alternating sign from check to check), so its
    residual settles near 1e-5; a case continued in FP64 has none."""

    def __init__(self, clock):
        super().__init__(clock)
        self.fp64 = set()

    def kinetic_energy(self, r):
        ke = 1.0 + np.exp(-r.steps / 2000.0)
        if r.rid in self.fp64:
            return ke
        return ke * (1.0 + 5e-6 * (1 if (r.steps // 1000) % 2 else -1))

    def refine(self, r):
        self.fp64.add(r.rid)
        return r.bytes_estimate


def test_fp64_continuation_reaches_the_tolerance_fp32_cannot():
    """Below the FP32 residual floor only the FP64 continuation converges: with
    switch_tol every case continues in FP64 and reaches 1e-8; without it none
    does before the step cap."""
    runs = {}
    for switch in (None, 1e-4):
        clk = FakeClock()
        eng = FloorEngine(clk)
        srv = Server(eng, hbm_budget=10 ** 6, check_every=1000, ke_tol=1e-8, max_steps=60000,
                     clock=clk, sleep=clk.sleep, switch_tol=switch)
        runs[switch] = srv.run(make_requests(5))
    assert runs[None]["converged"] == 0 and runs[None]["continued_fp64"] == 0
    assert runs[1e-4]["converged"] == 5 and runs[1e-4]["continued_fp64"] == 5
    assert 0 < runs[1e-4]["mean_fp64_steps"] < runs[1e-4]["mean_steps"]
