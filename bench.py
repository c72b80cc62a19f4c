"""Benchmark of the DUGKS step on one MI355X.

Workload: the Rayleigh-Taylor case, 256 x 1024 cells, periodic in x and
bounce-back walls in y, Allen-Cahn + D2Q9 (HCZ momentum closure), CD faces,
gravity, Atwood 0.5, Re 3000, seeded interface perturbation (synthetic input,
built on the host and uploaded once; the timed region starts with the state
resident in HBM).

A step is one full DUGKS time step of the case (phases 1-12 of the reference
loop).  The headline value is the FP32-bulk fused kernel's MLUPS; the plain
FP64 path (one kernel per phase) and the FP64 fused kernel are timed on the
same workload and reported beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MLUPS (FP64 plain vs FP32 fused) + %HBM peak; ensemble solves/s, p99 latency"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per lattice update: the state a cell must carry between
# steps, read once and written once.  AC+HCZ: 9 h~ + 9 f~ + prevPhiU/V + Phi of
# the previous step = 21 scalars.
S_ALG_HCZ = 21


def run_gpu(cfg, st, steps, warmup, path):
    from mpdugks_amd import solver
    s = solver.Solver(cfg, path=path)
    s.upload(st)
    s.step(warmup)
    s.synchronize()
    # HIP events on the solver's stream around the whole timed region (one
    # mpd_step call of `steps` back-to-back launches): kernel time per step
    s.set_kernel_timing(False)
    s.kernel_timing()
    t0 = time.perf_counter()
    s.step(steps)
    s.synchronize()
    wall = time.perf_counter() - t0
    avg_ms, nlaunch = s.kernel_timing()
    wall_free = wall
    tile = s.tile()
    red = s.reductions()
    out = s.download()
    s.close()
    return {"wall_s": wall, "wall_s_untimed": wall_free, "kernel_ms": avg_ms, "launches": nlaunch,
            "state": out, "sums": red, "tile": tile}


def ensemble_bench(ncases, steps, seed=0, solve_steps=20000):
    """Batched ensemble (one launch per step over all cases, HIP-graph replay)
    against the plain FP64 path run case by case; solves/s extrapolated to
    `solve_steps` steps per case from the measured per-step times."""
    from mpdugks_amd import abi, cases, solver
    specs = cases.ensemble_specs(ncases, seed=seed)
    members = []
    t0 = time.perf_counter()
    batch = solver.Batch()
    for sp in specs:
        cfg, st = cases.ensemble_case(sp)
        s = solver.Solver(cfg, path=abi.MPD_PATH_FUSED)
        s.upload(st)
        s.step(2)                      # two explicit steps prime the derived state
        batch.add(s)
        members.append(s)
    setup = time.perf_counter() - t0
    ncells = batch.size()[1]
    batch.step(30)                     # includes the tile-shape tuning (8 steps per shape)
    tile, tune_ms = batch.tile()
    batch.timing()
    t0 = time.perf_counter()
    batch.step(steps)
    wall = time.perf_counter() - t0
    ms_batch, _ = batch.timing()
    red = [m.reductions() for m in members[:8]]
    batch.close()
    for m in members:
        m.close()
    t_solve_batch = ms_batch * 1e-3 * solve_steps
    # plain FP64 path, one case at a time: per-step time for each mesh side
    per_side = {}
    for side in sorted(set(sp["side"] for sp in specs)):
        sp = dict(next(x for x in specs if x["side"] == side))
        cfg, st = cases.ensemble_case(sp, precision=abi.MPD_FP64)
        s = solver.Solver(cfg, path=abi.MPD_PATH_PHASES)
        s.upload(st)
        s.step(10)
        s.synchronize()
        n = max(20, steps // 10)
        t1 = time.perf_counter()
        s.step(n)
        s.synchronize()
        per_side[side] = (time.perf_counter() - t1) / n
        s.close()
    t_solve_plain = sum(per_side[sp["side"]] for sp in specs) * solve_steps
    return {"cases": ncases, "cells": ncells, "solve_steps": solve_steps,
            "batch_ms_per_step": ms_batch, "batch_mlups": ncells / (ms_batch * 1e-3) / 1e6,
            "solves_per_s_fp32_batched": ncases / t_solve_batch,
            "solves_per_s_fp64_plain": ncases / t_solve_plain,
            "speedup_vs_fp64_plain": t_solve_plain / t_solve_batch,
            "fp64_plain_ms_per_step_by_side": {str(k): v * 1e3 for k, v in per_side.items()},
            "tile": abi.TILE_NAMES[tile],
            "tuning_ms_per_step": {abi.TILE_NAMES[t + 1]: ms for t, ms in enumerate(tune_ms)},
            "setup_s": setup, "timed_steps": steps,
            "nonfinite_in_first_8": int(sum(r["nonfinite"] for r in red))}


def pp_bench(n=512, steps=1000, tr=0.85, seed=3, cmp_steps=200):
    """Pseudopotential liquid-vapour separation (D2Q9, Carnahan-Starling EOS):
    n x n periodic, rho0 = 0.1645 (1 + U(-0.01, 0.01)) per cell built on the
    device.  T/Tc = 0.85, the reference's shipped value: at 0.8 the reference
    algorithm itself diverges (non-finite by step 1000 on 128^2, FP64); at 0.85
    this start stays finite to step 2,500 on 512^2 (non-finite by 3,000, FP64 and
    FP32 alike), so the timed steps stay well inside that.
    FP32-bulk (FP32 populations, FP64 EOS and conserved density) against the
    plain FP64 phase path: MLUPS, mass drift, and the relative L2 distance of
    rho after `cmp_steps` steps from the same start."""
    from mpdugks_amd import abi, solver
    from mpdugks_amd.config import default_config
    out = {"workload": "pp_liquid_vapour_%dx%d" % (n, n), "tr": tr, "rho0": 0.1645, "noise": 0.01, "seed": seed}
    fields = {}
    for tag, prec in (("fp32_bulk", abi.MPD_FP32_BULK), ("fp64_plain", abi.MPD_FP64)):
        cfg = default_config(abi.MPD_MODEL_PP_D2Q9, nx=n, ny=n, pp_tr=tr)
        cfg.precision = prec
        s = solver.Solver(cfg)
        s.init_pp_noise(0.1645, amp=0.01, seed=seed)
        m0 = s.reductions()["sum_rho"]
        s.step(cmp_steps)
        fields[tag] = s.download().m[abi.M["rho"]]
        s.synchronize()
        k = steps if prec == abi.MPD_FP32_BULK else max(1, steps // 2)
        t0 = time.perf_counter()
        s.step(k)
        s.synchronize()
        dt = time.perf_counter() - t0
        r = s.reductions()
        s.close()
        out[tag] = {"steps": k, "ms_per_step": dt / k * 1e3, "mlups": n * n * k / dt / 1e6,
                    "mass_drift_rel": abs(r["sum_rho"] - m0) / m0, "nonfinite": r["nonfinite"]}
    a, b = fields["fp32_bulk"], fields["fp64_plain"]
    out["fp32_vs_fp64_rho_rel_l2_at_%d" % cmp_steps] = float(np.sqrt(((a - b) ** 2).sum()) / np.sqrt((b ** 2).sum()))
    out["speedup_fp32_vs_fp64_plain"] = out["fp32_bulk"]["mlups"] / out["fp64_plain"]["mlups"]
    out["d2v16"] = d2v16_bench(n)
    return out


def d2v16_bench(n=512, steps=400, cmp_steps=200, rho0=0.02, tr=0.8, seed=3):
    """The D2V16 velocity set on the same mesh: FP32-bulk (one launch per step)
    against the plain FP64 path, from rho0 = 0.02 (1 +- 0.1 %) at T/Tc = 0.8,
    the vapour side where the 16-velocity step grows slowest (about 1.002 per
    step at most), so the timed steps stay bounded."""
    from mpdugks_amd import abi, solver
    from mpdugks_amd.config import default_config
    out = {"workload": "pp_d2v16_%dx%d" % (n, n), "tr": tr, "rho0": rho0, "noise": 1e-3, "seed": seed}
    fields = {}
    for tag, prec in (("fp32_bulk", abi.MPD_FP32_BULK), ("fp64_plain", abi.MPD_FP64)):
        cfg = default_config(abi.MPD_MODEL_PP_D2V16, nx=n, ny=n, pp_tr=tr)
        cfg.precision = prec
        s = solver.Solver(cfg)
        s.init_pp_noise(rho0, amp=1e-3, seed=seed)
        s.step(cmp_steps)
        fields[tag] = s.download().m[abi.M["rho"]] - rho0
        s.synchronize()
        k = steps if prec == abi.MPD_FP32_BULK else max(1, steps // 4)
        t0 = time.perf_counter()
        s.step(k)
        s.synchronize()
        dt = time.perf_counter() - t0
        r = s.reductions()
        s.close()
        out[tag] = {"steps": k, "ms_per_step": dt / k * 1e3, "mlups": n * n * k / dt / 1e6,
                    "nonfinite": r["nonfinite"]}
    a, b = fields["fp32_bulk"], fields["fp64_plain"]
    out["fp32_vs_fp64_perturbation_rel_l2_at_%d" % cmp_steps] = float(np.sqrt(((a - b) ** 2).sum()) /
                                                                     np.sqrt((b ** 2).sum()))
    out["speedup_fp32_vs_fp64_plain"] = out["fp32_bulk"]["mlups"] / out["fp64_plain"]["mlups"]
    return out


def serving_bench(nreq, max_steps, budget_gb, glups, ke_tol=1e-5, seed=0):
    """Poisson stream of two-droplet coalescence requests (radii U(10,60), gap
    U{2..8}, mesh side 64-1024, density ratio 10, sigma 1e-2), FP32-bulk, each
    run to a kinetic-energy residual of `ke_tol` checked every 1000 steps (cap
    `max_steps`).  FP32-bulk's residual floor is 1e-6..3e-5 (the FP64 path
    reaches 1e-8), so the default stop is 1e-5.  The mean arrival rate puts the
    offered work (cells x expected steps) at ~80% of `glups`."""
    from mpdugks_amd import serve
    probe = serve.coalescence_requests(nreq, rate=1.0, seed=seed)
    mean_work = np.mean([r.cells * min(max_steps, serve.expected_steps(r)) for r in probe])
    rate = 0.8 * glups * 1e9 / mean_work
    reqs = serve.coalescence_requests(nreq, rate=rate, seed=seed)
    eng = serve.GpuEngine()
    srv = serve.Server(eng, hbm_budget=int(budget_gb * 2 ** 30), check_every=1000, ke_tol=ke_tol,
                       max_steps=max_steps)
    stats = srv.run(reqs)
    eng.close()
    stats.update({"requests": len(reqs), "arrival_rate_per_s": rate, "max_steps": max_steps,
                  "hbm_budget_gb": budget_gb, "ke_tol": ke_tol,
                  "achieved_glups": stats["cell_steps"] / stats["wall_s"] / 1e9})
    return stats


def copy_calibration(mib):
    """A plain device-to-device copy of `mib` MiB (reads and writes exactly that
    many bytes), run 5 times; its FETCH_SIZE / WRITE_SIZE per dispatch give the
    scale of the counters for the traffic figure."""
    import torch
    n = int(mib * (1 << 20)) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print("[bench] copy calibration: %d bytes per copy" % (4 * n), file=sys.stderr, flush=True)


def cpu_baseline(cfg, st, budget_s):
    """The CPU oracle (single thread, FP64) on the same case for ~budget_s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    o = oracle.Oracle(cfg)
    o.set_state(st.h, st.f, st.m)
    n = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        o.step(1)
        n += 1
    dt = time.perf_counter() - t0
    return cfg.nx * cfg.ny * n / dt / 1e6, n, dt


def cpu_reference(first, last, threads):
    """The reference solver (oracle/_ref/hcz_rt/ark, built here from the
    reference's own sources with the Rayleigh-Taylor mesh constants, 256x1024,
    bounce-back walls in y; the reference has no gravity) on `threads` OpenMP
    threads: steps first..last-1 timed between the harness's step-start
    timestamps (mesh reading and set-up excluded).  None when the binary is
    absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ref_run
    try:
        r = ref_run.time_window("hcz_rt", first, last, threads=threads, timeout=600)
    except Exception as e:  # a missing or failing reference binary only drops this field
        print("[bench] reference CPU solver not timed: %s" % e, file=sys.stderr, flush=True)
        return None
    if not r:
        return None
    n, sec = r
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count()
    return {"value": round(256 * 1024 * n / sec / 1e6, 3), "unit": "MLUPS", "cores": threads, "kind": "reference",
            "sample": "reference Ark OpenMP solver (g++ -O3 -fopenmp, its own sources), Rayleigh-Taylor mesh 256x1024, "
                      "AC+HCZ, bounce-back walls, no gravity; steps %d-%d (%.2f s) on %d threads (nproc %d, "
                      "affinity %d)" % (first, last, sec, threads, os.cpu_count() or 0, affinity)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-ref-steps", type=int, default=200,
                    help="steps of the reference solver timed (after 10 start-up steps)")
    ap.add_argument("--cpu-threads", type=int,
                    default=int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1),
                    help="OpenMP threads of the reference solver (default: OMP_NUM_THREADS, else nproc)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--ny", type=int, default=1024)
    ap.add_argument("--only", choices=["fp32_fused", "fp64_fused", "fp64_plain", "pp", "ensemble"], default=None,
                    help="run one variant only (for profiling); prints its numbers, not the headline line")
    ap.add_argument("--ensemble-cases", type=int, default=256)
    ap.add_argument("--ensemble-steps", type=int, default=400)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=24)
    ap.add_argument("--serve-max-steps", type=int, default=400000)
    ap.add_argument("--serve-ke-tol", type=float, default=1e-5)
    ap.add_argument("--serve-budget-gb", type=float, default=2.0)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--pp-steps", type=int, default=1000)
    ap.add_argument("--no-pp", action="store_true")
    ap.add_argument("--copy-mb", type=float, default=0.0,
                    help="also run a plain device-to-device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU only (the DUGKS case runs on a single "
              "MI355X); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from mpdugks_amd import abi, cases
    nx, ny = args.nx, args.ny
    N = nx * ny
    if args.copy_mb > 0:
        copy_calibration(args.copy_mb)
    if args.only == "pp":
        print(json.dumps(pp_bench(steps=args.pp_steps)))
        return
    if args.only == "ensemble":
        print(json.dumps(ensemble_bench(args.ensemble_cases, args.ensemble_steps)))
        return
    results = {}
    variants = (("fp32_fused", abi.MPD_FP32_BULK, abi.MPD_PATH_FUSED),
                ("fp64_fused", abi.MPD_FP64, abi.MPD_PATH_FUSED),
                ("fp64_plain", abi.MPD_FP64, abi.MPD_PATH_PHASES))
    # The host-side cases are built first and the side benchmarks (ensemble,
    # serving, pseudopotential) run before the timed Rayleigh-Taylor regions,
    # so that those start on a GPU that is already working rather than one
    # coming out of idle: a 20-step region measured 15.7-16.1 us per step on
    # the step kernel's events after upload or 0.5 s of idle and 14.9-15.1 us
    # after sustained work, the steady-state figure.
    prepared = {tag: cases.rayleigh_taylor(nx=nx, ny=ny, precision=prec) for tag, prec, _ in variants
                if not args.only or tag == args.only}
    side = {}
    if not args.only:
        if not args.no_ensemble:
            side["ensemble"] = ensemble_bench(args.ensemble_cases, args.ensemble_steps)
        if not args.no_serve:
            glups = side.get("ensemble", {}).get("batch_mlups", 10000.0) / 1e3
            side["serving"] = serving_bench(args.serve_requests, args.serve_max_steps, args.serve_budget_gb,
                                            glups, ke_tol=args.serve_ke_tol)
        if not args.no_pp:
            side["pseudopotential"] = pp_bench(steps=args.pp_steps)
    for tag, prec, path in variants:
        if tag not in prepared:
            continue
        cfg, st = prepared.pop(tag)
        steps = args.steps if tag != "fp64_plain" else max(1, args.steps // 4)
        r = run_gpu(cfg, st, steps, args.warmup, path)
        r["steps"] = steps
        r["mlups"] = N * steps / r["wall_s_untimed"] / 1e6
        r["kernel_mlups"] = N / (r["kernel_ms"] * 1e-3) / 1e6 if r["kernel_ms"] > 0 else 0.0
        results[tag] = r
        print("[bench] %s: %d steps, %.4f ms/step wall, kernel %.4f ms, %.1f MLUPS"
              % (tag, steps, r["wall_s_untimed"] / steps * 1e3, r["kernel_ms"], r["mlups"]), file=sys.stderr,
              flush=True)

    if args.only:
        r = results[args.only]
        print(json.dumps({"variant": args.only, "steps": r["steps"], "mlups": r["mlups"],
                          "kernel_ms": r["kernel_ms"], "ms_per_step": r["wall_s_untimed"] / r["steps"] * 1e3}))
        return
    head = results["fp32_fused"]
    ref64 = results["fp64_fused"]["state"]
    g32 = head["state"]
    # FP32-bulk error against the FP64 fused run after the same steps (both
    # ran warmup + steps)
    def rel(a, b):
        return float(np.sqrt(((a - b) ** 2).sum()) / max(np.sqrt((b ** 2).sum()), 1e-300))
    err = {k: rel(g32.m[abi.M[k]], ref64.m[abi.M[k]]) for k in ("rho", "phi", "p")}
    # the case's own sensitivity: the FP64 fused run from the start state with
    # every population multiplied by 1 + 6e-8 U(-1, 1) (FP32 round-off size),
    # after the same number of steps
    from mpdugks_amd import solver
    cfg64, st64 = cases.rayleigh_taylor(nx=nx, ny=ny, precision=abi.MPD_FP64)
    rng = np.random.default_rng(1)
    st64.h = st64.h * (1 + 6e-8 * rng.uniform(-1, 1, st64.h.shape))
    st64.f = st64.f * (1 + 6e-8 * rng.uniform(-1, 1, st64.f.shape))
    sp = solver.Solver(cfg64, path=abi.MPD_PATH_FUSED)
    sp.upload(st64)
    sp.step(args.warmup + results["fp64_fused"]["steps"])
    pert = sp.download()
    sp.close()
    err_floor = {k: rel(pert.m[abi.M[k]], ref64.m[abi.M[k]]) for k in ("rho", "phi", "p")}

    bytes_alg = 2 * S_ALG_HCZ * 4  # FP32-bulk
    kernel = "k_fused32<AC_HCZ,BB walls,%s tiles>" % abi.TILE_NAMES[head["tile"]]
    traffic = None
    try:  # HBM bytes per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE runs
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as fh:
            rec = json.load(fh).get("%s rayleigh_taylor_%dx%d" % (kernel, nx, ny))
        traffic = rec["traffic_bytes"] if rec else None
    except (OSError, ValueError):
        traffic = None
    achieved = bytes_alg * N / (head["kernel_ms"] * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": round(head["mlups"], 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": head["steps"],
        "warmup": args.warmup,
        "ms_per_step": head["wall_s_untimed"] / head["steps"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "rayleigh_taylor_%dx%d" % (nx, ny), "cells": N,
                   "physics": "Allen-Cahn + D2Q9 HCZ, CD faces, periodic x, bounce-back walls y, gravity",
                   "atwood": 0.5, "reynolds": 3000, "precision": "fp32-bulk (fused kernel)",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": kernel,
                     "alg_bytes_per_cell": bytes_alg, "kernel_ms": head["kernel_ms"]},
        "fp64_plain_mlups": round(results["fp64_plain"]["mlups"], 2),
        "fp64_fused_mlups": round(results["fp64_fused"]["mlups"], 2),
        "fp32_fused_kernel_mlups": round(head["kernel_mlups"], 2),
        "fp32_vs_fp64_rel_l2": err,
        "fp64_perturbed_vs_fp64_rel_l2": err_floor,
    }
    if "ensemble" in side:
        ens = side["ensemble"]
        line["ensemble"] = ens
        print("[bench] ensemble: %d cases, %d cells, batch %.3f ms/step, %.2f solves/s (FP32 batched) vs "
              "%.3f (FP64 plain), x%.1f" % (ens["cases"], ens["cells"], ens["batch_ms_per_step"],
                                             ens["solves_per_s_fp32_batched"], ens["solves_per_s_fp64_plain"],
                                             ens["speedup_vs_fp64_plain"]), file=sys.stderr, flush=True)
    if "serving" in side:
        sv = side["serving"]
        line["serving"] = sv
        conv = ("p50 %.3f s, p99 %.3f s" % (sv["p50_latency_converged_s"], sv["p99_latency_converged_s"])
                if sv["converged"] else "none")
        print("[bench] serving: %d done, %d converged to KE residual %.0e, %.2f solves/s, p50 %.3f s, "
              "p99 %.3f s (converged requests: %s), busy %.2f, mean %.0f steps"
              % (sv["completed"], sv["converged"], sv["ke_tol"], sv["solves_per_s"], sv["p50_latency_s"],
                 sv["p99_latency_s"], conv, sv["gpu_busy_frac"], sv["mean_steps"]), file=sys.stderr, flush=True)
    if "pseudopotential" in side:
        pp = side["pseudopotential"]
        line["pseudopotential"] = pp
        print("[bench] pseudopotential 512^2: FP32-bulk %.1f MLUPS, FP64 plain %.1f MLUPS, mass drift %.1e, "
              "rho rel L2 %.2e" % (pp["fp32_bulk"]["mlups"], pp["fp64_plain"]["mlups"],
                                   pp["fp32_bulk"]["mass_drift_rel"], pp["fp32_vs_fp64_rho_rel_l2_at_200"]),
              file=sys.stderr, flush=True)
    if not args.no_cpu:
        cfg, st = cases.rayleigh_taylor(nx=nx, ny=ny, precision=abi.MPD_FP64)
        mlups, nsteps, dt = cpu_baseline(cfg, st, args.cpu_seconds)
        port = {"value": round(mlups, 3), "unit": "MLUPS", "cores": 1, "kind": "port",
                "sample": "CPU FP64 oracle, same RT %dx%d case, first %d steps (%.1f s)" % (nx, ny, nsteps, dt)}
        ref = cpu_reference(10, 10 + args.cpu_ref_steps, args.cpu_threads) if (nx, ny) == (256, 1024) else None
        if ref:
            line["cpu_baseline"] = ref
            line["cpu_port"] = port
            print("[bench] reference CPU solver, RT 256x1024: %.3f MLUPS on %d threads"
                  % (ref["value"], ref["cores"]), file=sys.stderr, flush=True)
        else:
            line["cpu_baseline"] = port
    print(json.dumps(line))


if __name__ == "__main__":
    main()
