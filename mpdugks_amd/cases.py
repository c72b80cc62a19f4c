"""Benchmark and test cases (synthetic, seeded): the Rayleigh-Taylor instability,
the static droplet/bubble ensemble and two-droplet coalescence."""
import numpy as np

from . import abi, solver
from .config import default_config


def rayleigh_taylor(nx=256, ny=1024, precision=abi.MPD_FP32_BULK, seed=0, width=4.0, sigma=5e-5,
                    atwood=0.5, reynolds=3000.0, sqrt_gw=0.04):
    """Rayleigh-Taylor instability: periodic in x, bounce-back walls in y, heavy
    phase (phi = 1) above an interface at mid-height perturbed by
    0.1*W*cos(2*pi*x/W) plus seeded uniform noise of amplitude 1e-3*W (W = nx).
    Atwood 0.5 -> rho_heavy/rho_light = 3; Re = sqrt(g W) W / nu = 3000 with
    sqrt(g W) = 0.04 -> g = 6.25e-6, nu = 3.413e-3; gravity F_y = -g (rho - rho_mean)."""
    rho_light = 1.0
    rho_heavy = rho_light * (1 + atwood) / (1 - atwood)
    W = float(nx)
    g = sqrt_gw ** 2 / W
    nu = sqrt_gw * W / reynolds
    cfg = default_config(abi.MPD_MODEL_AC_HCZ, nx=nx, ny=ny)
    cfg.bc_y = abi.MPD_BC_WALL_BB
    cfg.precision = precision
    cfg.rho_l, cfg.rho_v = rho_heavy, rho_light
    cfg.nu_l = cfg.nu_v = nu
    cfg.width = width
    cfg.sigma = sigma
    cfg.gy = g
    cfg.rho_ref = 0.5 * (rho_heavy + rho_light)
    st = solver.ac_interface(cfg, rayleigh_taylor_y0(cfg, seed), sign=+1)
    return cfg, st


def rayleigh_taylor_y0(cfg, seed=0):
    """Interface heights of the Rayleigh-Taylor start: mid-height + 0.1 W cos(2 pi x / W)
    + U(-1, 1) * 1e-3 W per column (W = nx, seeded)."""
    W = float(cfg.nx)
    rng = np.random.default_rng(seed)
    x = (np.arange(cfg.nx) + 0.5) * cfg.dx
    return 0.5 * cfg.ny * cfg.dx + 0.1 * W * np.cos(2 * np.pi * x / W) + rng.uniform(-1, 1, cfg.nx) * 1e-3 * W


ENSEMBLE_SIDES = (64, 96, 128, 192, 256)


def ensemble_specs(n=256, seed=0, sides=ENSEMBLE_SIDES):
    """Parameters of the batched ensemble: n static droplet/bubble cases, mesh
    side uniform over `sides`, radius U(0.15, 0.35)*side, density ratio
    log-uniform in [2, 1000], surface tension log-uniform in [1e-4, 1e-2],
    droplet or bubble by a coin flip; all seeded."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(n):
        side = int(rng.choice(sides))
        specs.append({
            "side": side,
            "radius": float(rng.uniform(0.15, 0.35) * side),
            "ratio": float(np.exp(rng.uniform(np.log(2.0), np.log(1000.0)))),
            "sigma": float(np.exp(rng.uniform(np.log(1e-4), np.log(1e-2)))),
            "sign": int(rng.choice([-1, 1])),
        })
    return specs


def ensemble_case(spec, precision=abi.MPD_FP32_BULK, model=abi.MPD_MODEL_AC_HCZ, width=4.0):
    """One ensemble member: periodic square mesh, AC + D2Q9, drop of the spec."""
    side = spec["side"]
    cfg = default_config(model, nx=side, ny=side)
    cfg.precision = precision
    cfg.rho_v = 1.0
    cfg.rho_l = spec["ratio"]
    cfg.sigma = spec["sigma"]
    cfg.width = width
    st = solver.ac_drop(cfg, radius=spec["radius"], sign=spec["sign"])
    return cfg, st


def coalescence_case(r1, r2, gap, precision=abi.MPD_FP32_BULK, width=4.0, sides=(64, 96, 128, 192, 256, 384, 512, 768, 1024),
                     ratio=10.0, sigma=1e-3):
    """Two droplets side by side (radii r1, r2, gap cells between interfaces);
    mesh side = smallest of `sides` that holds 2(r1+r2) + gap + 4W."""
    need = 2 * (r1 + r2) + gap + 4 * width
    side = next(s for s in sides if s >= need)
    cfg = default_config(abi.MPD_MODEL_AC_HCZ, nx=side, ny=side)
    cfg.precision = precision
    cfg.rho_v, cfg.rho_l = 1.0, ratio
    cfg.sigma = sigma
    cfg.width = width
    (x1, c, _), (x2, _, _) = coalescence_drops(r1, r2, gap, cfg)
    a = solver.ac_drop(cfg, cx=x1, cy=c, radius=r1, sign=-1)
    b = solver.ac_drop(cfg, cx=x2, cy=c, radius=r2, sign=-1)
    # union of the two droplets: phi = max of the two profiles, then equilibrium
    phi = np.maximum(a.m[abi.M["phi"]], b.m[abi.M["phi"]])
    st = a if a.m[abi.M["phi"]].sum() >= 0 else b
    sel = b.m[abi.M["phi"]] > a.m[abi.M["phi"]]
    for arr_a, arr_b in ((a.h, b.h), (a.f, b.f), (a.m, b.m)):
        arr_a[:, sel] = arr_b[:, sel]
    assert np.allclose(a.m[abi.M["phi"]], phi)
    return cfg, a


def coalescence_config(r1, r2, gap, precision=abi.MPD_FP32_BULK, width=4.0,
                       sides=(64, 96, 128, 192, 256, 384, 512, 768, 1024), ratio=10.0, sigma=1e-3):
    """The configuration of coalescence_case without building the state (for a
    start built on the device): (cfg, drops)."""
    need = 2 * (r1 + r2) + gap + 4 * width
    side = next(s for s in sides if s >= need)
    cfg = default_config(abi.MPD_MODEL_AC_HCZ, nx=side, ny=side)
    cfg.precision = precision
    cfg.rho_v, cfg.rho_l = 1.0, ratio
    cfg.sigma = sigma
    cfg.width = width
    return cfg, coalescence_drops(r1, r2, gap, cfg)


def coalescence_drops(r1, r2, gap, cfg):
    """Centres and radii of the two droplets, side by side at mid-height with
    `gap` cells between their interfaces: [(cx, cy, r), (cx, cy, r)]."""
    c = 0.5 * cfg.nx * cfg.dx
    x1 = c - 0.5 * (2 * r1 + 2 * r2 + gap) + r1
    x2 = x1 + r1 + gap + r2
    return [(x1, c, r1), (x2, c, r2)]
