"""Run-time configuration mirroring the reference's compile-time constants.

Defaults are the shipped values of Ark/Constant/ZeroReference.h (mesh, RT,
Mu0, Tau0) and Ark/Constant/ZeroNamespace.h (CFL, the PhaseFieldAC and
PseudoPotentialSC namespaces), evaluated with the same expressions so the
doubles match bit for bit.
"""
import math

from .abi import (MpdConfig, MPD_MODEL_AC_HCZ, MPD_MODEL_PP_D2Q9, MPD_MODEL_PP_D2V16,
                  MPD_FLUX_CD, MPD_BC_PERIODIC, MPD_FP64)


def default_config(model=MPD_MODEL_AC_HCZ, nx=128, ny=128, **overrides):
    ch_length = 128.0                       # ZeroReference.h:18
    min_l = ch_length / 128                 # ZeroReference.h:20
    t0, r0 = 1.0 / 3, 1.0                   # ZeroReference.h:70-72
    rt = r0 * t0
    mu0, rho0 = 0.1, 1.0                    # ZeroReference.h:78,92
    nu0 = mu0 / rho0
    tau0 = nu0 / rt                         # ZeroReference.h:100
    maxu = math.sqrt(3.0 * rt)              # D2Q9.h:16
    cfl = 0.5                               # ZeroNamespace.h:10
    dt = cfl * min_l / maxu if cfl > 0.0 else 1.0e-4
    cn = 5 / ch_length                      # ZeroNamespace.h:65
    cfg = MpdConfig()
    cfg.nx, cfg.ny = nx, ny
    cfg.model = model
    cfg.flux = MPD_FLUX_CD
    cfg.bc_y = MPD_BC_PERIODIC
    cfg.precision = MPD_FP64
    cfg.dx = min_l
    cfg.dt = dt
    cfg.rt = rt
    cfg.phi_l, cfg.phi_v = 1.0, 0.0         # ZeroNamespace.h:79
    cfg.rho_l, cfg.rho_v = 10.0, 1.0        # ZeroNamespace.h:81
    cfg.nu_l, cfg.nu_v = nu0, nu0           # ZeroNamespace.h:89
    cfg.mobility = 0.1                      # ZeroNamespace.h:63
    cfg.width = ch_length * cn              # ZeroNamespace.h:69
    cfg.sigma = 1e-2                        # ZeroNamespace.h:71
    cfg.tau0 = tau0
    cfg.gx = cfg.gy = 0.0                   # ZeroNamespace.h:95
    cfg.rho_ref = 0.0
    cfg.pp_tr = 0.85                        # ZeroNamespace.h:142
    cfg.pp_k = 1.0                          # ZeroNamespace.h:140
    cfg.pp_tau = tau0                       # Preprocess.cpp:376,397
    if model == MPD_MODEL_PP_D2V16:
        # MaxU = 2 MaSpan sqrt(2) with MaSpan = 1.2 (Ark/Lib/D2V16.h:16-20)
        cfg.dt = cfl * min_l / (2 * 1.2 * math.sqrt(2.0))
    for k, v in overrides.items():
        setattr(cfg, k, v)
    return cfg
