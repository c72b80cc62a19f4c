"""Shared helpers of the test suite (cases, digests, oracle runs)."""
import hashlib

import numpy as np

from mpdugks_amd import abi, solver
from mpdugks_amd.config import default_config


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def make_case(model="ac_hcz", nx=128, ny=128, flux=abi.MPD_FLUX_CD, walls=False, sign=+1,
              radius=None, precision=abi.MPD_FP64, nee=False, **over):
    m = abi.MODEL_NAMES[model]
    cfg = default_config(m, nx=nx, ny=ny, **over)
    cfg.flux = flux
    cfg.bc_y = (abi.MPD_BC_WALL_NEE if nee else abi.MPD_BC_WALL_BB) if walls else abi.MPD_BC_PERIODIC
    cfg.precision = precision
    if m == abi.MPD_MODEL_PP_D2Q9:
        st = solver.scmp_drop(cfg, radius=radius if radius is not None else 0.25 * nx)
    else:
        st = solver.ac_drop(cfg, radius=radius if radius is not None else 0.36 * nx, sign=sign)
    return cfg, st


def oracle_run(cfg, st, steps):
    """Run the CPU oracle from host state `st`; return states after each count in `steps`."""
    from oracle import Oracle
    o = Oracle(cfg)
    o.set_state(st.h if st.h is not None else np.zeros_like(st.f), st.f, st.m, st.ffx, st.ffy)
    out, done = {}, 0
    for k in steps:
        o.step(k - done)
        done = k
        h, f, m, ffx, ffy = o.get_state()
        out[k] = solver.State(cfg, h if abi.has_h(cfg.model) else None, f, m, ffx, ffy)
    return out


def max_rel(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
