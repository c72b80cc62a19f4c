"""Serving loop: simulation requests kept resident on the GPU in one batch.

Requests arrive over time (seeded Poisson process); each is admitted when its
arrival time has passed and its device memory fits the HBM budget, primed with
the two explicit steps the derived state needs, and added to the batch
(libmpdugks_hip's mpd_batch_*: one kernel launch per step over all resident
cases).  Every `check_every` steps each case's kinetic-energy residual
|KE^n - KE^(n-check_every)| / KE^n is evaluated from FP64 device sums; a case
retires when it falls below `ke_tol` or reaches `max_steps`.  Latency is
completion time minus arrival time (wall clock).

With FP32-bulk populations the residual of a coalesced drop settles on a
rounding floor of about 1e-6..3e-5 (the kinetic energy of the spurious
currents, 1e-6 of the state's scale, carries FP32 rounding); the FP64 path
goes below 1e-8.  A tolerance under that floor is reached only by chance, so
with `switch_tol` set a request that reaches it in the FP32 batch continues as
an FP64 handle from its downloaded state (stepped on its own stream beside the
batch) until `ke_tol`: the batch does the long relaxation, FP64 the last
decades.

The reference runs one case per process with a density residual stop
(Ark/Cpp/DUGKS2DSolver.cpp:807-820); this loop and its KE residual are new.
"""
import collections
import time

import numpy as np


class Request:
    __slots__ = ("rid", "arrival", "make", "bytes_estimate", "admitted", "done", "steps", "converged",
                 "ke_prev", "handle", "cells", "final", "params", "fp64_from")

    def __init__(self, rid, arrival, make, bytes_estimate, cells, params=None):
        self.rid = rid
        self.arrival = arrival
        self.make = make              # () -> (cfg, state), or (cfg, drops) for a start built on the GPU
        self.bytes_estimate = bytes_estimate
        self.cells = cells
        self.admitted = None
        self.done = None
        self.steps = 0
        self.converged = False
        self.ke_prev = None
        self.handle = None
        self.final = None             # state at retirement (GpuEngine(keep_final=True))
        self.params = params or {}
        self.fp64_from = None         # step at which the case continued in FP64


class GpuEngine:
    """Runs cases through the library's batch API."""

    def __init__(self, keep_final=False):
        from . import solver
        self._solver = solver
        self.batch = solver.Batch()
        self.keep_final = keep_final
        self.fp64 = []                # FP64 handles, each stepped on its own stream

    def admit(self, req):
        from . import abi
        cfg, st = req.make()
        s = self._solver.Solver(cfg, path=abi.MPD_PATH_FUSED)
        try:
            if isinstance(st, list):
                s.init_drops(st, sign=-1)   # droplets built on the device: no host state, no upload
            else:
                s.upload(st)
            s.step(2)
            self.batch.add(s)
        except Exception:
            s.close()
            raise
        req.handle = s
        req.steps = 2
        return s.device_bytes()

    def step(self, n):
        # all queued before any wait: the batch and the FP64 handles run side by side
        if self.batch.size()[0]:
            self.batch.step(n)
        for h in self.fp64:
            h.step(n)

    def refine(self, req):
        """Continue `req` as an FP64 handle from its current state.  Returns its
        device bytes so far: the phase scratch and the derived planes are
        allocated at its first step."""
        from . import abi
        st = req.handle.download()
        cfg = type(st.cfg).from_buffer_copy(st.cfg)
        cfg.precision = abi.MPD_FP64
        self.batch.remove(req.handle)
        req.handle.close()
        s = self._solver.Solver(cfg, path=abi.MPD_PATH_FUSED)
        s.upload(self._solver.State(cfg, st.h, st.f, st.m))
        self.fp64.append(s)
        req.handle = s
        return s.device_bytes()

    def kinetic_energy(self, req):
        return req.handle.reductions()["kinetic_energy"]

    def retire(self, req):
        if self.keep_final:
            req.final = req.handle.download()
        if req.fp64_from is not None:
            self.fp64.remove(req.handle)
        else:
            self.batch.remove(req.handle)
        req.handle.close()
        req.handle = None

    def close(self):
        for h in self.fp64:
            h.close()
        self.fp64 = []
        self.batch.close()


class Server:
    def __init__(self, engine, hbm_budget, check_every=1000, ke_tol=1e-8, max_steps=10 ** 6,
                 clock=time.perf_counter, sleep=time.sleep, switch_tol=None):
        self.engine = engine
        self.switch_tol = switch_tol  # FP32 residual at which a case continues in FP64 (None: never)
        self.budget = hbm_budget
        self.check_every = check_every
        self.ke_tol = ke_tol
        self.max_steps = max_steps
        self.clock = clock
        self.sleep = sleep
        self.used = 0
        self.rejected = []

    def run(self, requests):
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        active, finished = [], []
        busy = 0.0
        max_active, queued = 0, set()
        t0 = self.clock()
        while pending or active:
            now = self.clock() - t0
            # admission in arrival order while the HBM budget allows
            while pending and pending[0].arrival <= now:
                r = pending[0]
                if r.bytes_estimate > self.budget:
                    self.rejected.append(pending.popleft())   # can never fit
                    continue
                if self.used + r.bytes_estimate > self.budget:
                    queued.add(r.rid)     # arrived, waits for HBM budget
                    break
                pending.popleft()
                r.admitted = now
                r.bytes_estimate = max(r.bytes_estimate, self.engine.admit(r))
                self.used += r.bytes_estimate
                active.append(r)
            max_active = max(max_active, len(active))
            if not active:
                if pending:
                    self.sleep(max(0.0, min(pending[0].arrival - now, 0.01)))
                continue
            t1 = self.clock()
            self.engine.step(self.check_every)
            busy += self.clock() - t1
            now = self.clock() - t0
            still = []
            for r in active:
                r.steps += self.check_every
                ke = self.engine.kinetic_energy(r)
                res = abs(ke - r.ke_prev) / ke if (r.ke_prev is not None and ke > 0) else np.inf
                r.ke_prev = ke
                if res < self.ke_tol or r.steps >= self.max_steps:
                    r.converged = res < self.ke_tol
                    r.done = now
                    self.engine.retire(r)
                    self.used -= r.bytes_estimate
                    finished.append(r)
                    continue
                if self.switch_tol is not None and r.fp64_from is None and res < self.switch_tol:
                    # past what FP32 resolves: the rest in FP64, residual history restarted
                    # the FP32 estimate covers the FP64 handle's full footprint (its
                    # phase scratch and derived planes appear only at its first step)
                    self.used -= r.bytes_estimate
                    r.bytes_estimate = max(r.bytes_estimate, self.engine.refine(r))
                    self.used += r.bytes_estimate
                    r.fp64_from = r.steps
                    r.ke_prev = None
                still.append(r)
            active = still
        wall = self.clock() - t0
        lat = np.array([r.done - r.arrival for r in finished]) if finished else np.array([0.0])
        # the requests that met the tolerance (the others stopped at the step cap)
        conv = np.array([r.done - r.arrival for r in finished if r.converged])
        pct = lambda a, q: float(np.percentile(a, q)) if a.size else None
        return {
            "completed": len(finished),
            "rejected": len(self.rejected),
            "converged": int(sum(r.converged for r in finished)),
            "wall_s": wall,
            "solves_per_s": len(finished) / wall if wall > 0 else 0.0,
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "p50_latency_converged_s": pct(conv, 50),
            "p99_latency_converged_s": pct(conv, 99),
            "gpu_busy_frac": busy / wall if wall > 0 else 0.0,
            "cell_steps": int(sum(r.cells * r.steps for r in finished)),
            "max_active": max_active,
            "queued_for_budget": len(queued),
            "mean_steps": float(np.mean([r.steps for r in finished])) if finished else 0.0,
            "continued_fp64": int(sum(r.fp64_from is not None for r in finished)),
            "mean_fp64_steps": float(np.mean([r.steps - r.fp64_from for r in finished if r.fp64_from is not None]))
            if any(r.fp64_from is not None for r in finished) else 0.0,
        }


# steps to a KE residual of 1e-5 per unit of r1^2 + r2^2 (FP32-bulk, sigma 1e-2,
# nu = M = 0.1: 23k steps for r = 10, 71k for r = 20)
STEPS_PER_R2 = 100.0


def coalescence_requests(n, rate, seed=0, max_side=1024, sigma=1e-2, radii=(10.0, 60.0)):
    """n two-droplet requests with Poisson arrivals of mean rate `rate` (1/s):
    radii U(10, 60) cells, gap U{2..8}; mesh from cases.coalescence_config
    (density ratio 10, surface tension `sigma`), the droplets built on the GPU."""
    from . import cases
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for k in range(n):
        t += rng.exponential(1.0 / rate)
        r1, r2 = rng.uniform(radii[0], radii[1], 2)
        gap = int(rng.integers(2, 9))
        need = 2 * (r1 + r2) + gap + 16
        side = next(s for s in (64, 96, 128, 192, 256, 384, 512, 768, 1024) if s >= need)
        if side > max_side:
            continue
        make = (lambda a=r1, b=r2, g=gap: cases.coalescence_config(a, b, g, sigma=sigma))
        # device bytes: FP64 explicit state + phase scratch + two FP32 derived copies
        cells = side * side
        reqs.append(Request(k, t, make, bytes_estimate=cells * 1100, cells=cells,
                            params={"r1": float(r1), "r2": float(r2), "gap": gap, "side": side, "sigma": sigma}))
    return reqs


def expected_steps(req):
    """Rough step count of a request to a KE residual of 1e-5 (pacing only)."""
    return STEPS_PER_R2 * (req.params["r1"] ** 2 + req.params["r2"] ** 2)
