"""Benchmark and test cases named by BASELINE configs (synthetic, seeded)."""
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
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) * cfg.dx
    y0 = 0.5 * ny * cfg.dx + 0.1 * W * np.cos(2 * np.pi * x / W) + rng.uniform(-1, 1, nx) * 1e-3 * W
    st = solver.ac_interface(cfg, y0, sign=+1)
    return cfg, st
