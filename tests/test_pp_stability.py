"""Linear (von Neumann) stability of the pseudopotential step, D2Q9 and D2V16.

The oracle's step is linearised about a uniform state at rest by central
differences: for a Fourier mode exp(i k.x) the per-cell state perturbation
(populations, U and V -- they carry the previous step's force, so they are
state --, the cell force and the x- and y-face forces; rho is the sum of the
populations) maps to a new one through a real 2n x 2n matrix acting on the cos
and sin parts; its eigenvalues are the amplification factors of that mode per
step.  Modes run along x on a 32 x 4 mesh (k = 2 pi m / 32, m = 1..15) and
along the diagonal on a 16 x 16 mesh (m = 1..7); the Nyquist mode is left out
(its sine part vanishes at the face positions).

The machinery is checked against a closed form first: the slowest-decaying
mode of D2Q9PP at long wavelength is the shear wave, exp(-nu k^2 dt), nu = tau RT.

Findings, at T/Tc = 0.8 (Carnahan-Starling, a = 1, b = 4), tau = 0.3:
  * D2Q9PP is stable about both coexisting phases (vapour 0.02168, liquid
    0.30733): max |lambda| = 0.99253 over all modes.
  * D2V16 (c = 1.2, eta = 3.6, RT = 1/3) is unstable about both: max |lambda|
    = 1.0022 (vapour), 1.0180 (liquid); and for tau 0.3 .. 2, CFL 0.5 and 0.25
    alike (1.002 .. 1.035).
  * At rho = 0.33186, where dp/drho = RT so that the force vanishes to linear
    order, D2V16 still amplifies diagonal modes by 1.029 per step (k ~ 1.2):
    the 16-velocity DUGKS step itself is unstable at intermediate
    wavenumbers, independent of the force (its axis-2c rest weights are
    negative, -0.36 at RT = 1/3); D2Q9PP there is stable (max 0.998).
A D2V16 liquid-vapour run therefore cannot stay bounded however it is
implemented; the liquid-vapour workload runs on D2Q9PP.
"""
import math

import numpy as np
import pytest

from mpdugks_amd import abi, solver
from mpdugks_amd.config import default_config

from test_d2v16 import _cs_eos, _feq9

RHO_V, RHO_L = 0.021682, 0.30733     # Maxwell construction of the CS EOS at T/Tc = 0.8
RHO_FREE = 0.3318551377038003        # dp/drho = RT at T/Tc = 0.8


def amplification(model, rho0, modes, nx=32, ny=4, diag=False, tr=0.8, eps=1e-8):
    """[(k, max |eigenvalue|)] of the linearised oracle step about uniform rho0."""
    import oracle
    if diag:
        ny = nx
    cfg = default_config(model, nx=nx, ny=ny, pp_tr=tr)
    N = nx * ny
    f0 = oracle.d2v16_feq(rho0, 0, 0, cfg.rt) if model == abi.MPD_MODEL_PP_D2V16 else _feq9(cfg, rho0, 0, 0)
    st = solver.State(cfg)
    st.f[:] = f0[:, None]
    st.m[abi.M["rho"]] = rho0
    st.m[abi.M["tau"]] = cfg.pp_tau
    st.m[abi.M["p"]] = _cs_eos(rho0, tr)
    Q = st.f.shape[0]
    ncomp = Q + 8
    ii, jj = np.arange(N) % nx, np.arange(N) // nx
    cell = ii + 0.5 + ((jj + 0.5) if diag else 0)
    xface = ii + ((jj + 0.5) if diag else 0)
    yface = ii + 0.5 + (jj if diag else 0)
    pos = [cell] * (Q + 4) + [xface] * 2 + [yface] * 2
    mi = [abi.M[k] for k in ("u", "v", "fx", "fy")]

    def flat(f, m, ffx, ffy):
        return np.concatenate([f, m[mi], ffx[:N][None], ffy[:N][None], ffx[N:][None], ffy[N:][None]])

    base = flat(st.f, st.m, st.ffx, st.ffy)
    o = oracle.Oracle(cfg)

    def step(d):
        f = st.f + d[:Q]
        m = st.m.copy()
        m[abi.M["rho"]] = f.sum(0)
        for c, k in enumerate(mi):
            m[k] += d[Q + c]
        ffx, ffy = st.ffx.copy(), st.ffy.copy()
        ffx[:N] += d[Q + 4]; ffy[:N] += d[Q + 5]; ffx[N:] += d[Q + 6]; ffy[N:] += d[Q + 7]
        o.set_state(np.zeros_like(f), f, m, ffx, ffy)
        o.step(1)
        _, f2, m2, fx2, fy2 = o.get_state()
        return flat(f2, m2, fx2, fy2) - base

    out = []
    for mm in modes:
        k = 2 * math.pi * mm / nx
        cs = [(np.cos(k * p), np.sin(k * p)) for p in pos]
        G = np.zeros((2 * ncomp, 2 * ncomp))
        for c in range(ncomp):
            for b in range(2):
                d = np.zeros((ncomp, N))
                d[c] = eps * cs[c][b]
                r = (step(d) - step(-d)) / (2 * eps)
                for c2 in range(ncomp):
                    G[2 * c2, 2 * c + b] = (r[c2] * cs[c2][0]).sum() * 2 / N
                    G[2 * c2 + 1, 2 * c + b] = (r[c2] * cs[c2][1]).sum() * 2 / N
        out.append((k, float(np.abs(np.linalg.eigvals(G)).max())))
    return cfg, out


def max_amplification(model, rho0):
    _, a = amplification(model, rho0, range(1, 16))
    _, b = amplification(model, rho0, range(1, 8), nx=16, diag=True)
    return max(l for _, l in a + b)


def test_linearisation_reproduces_shear_decay(oracle_mod):
    cfg, out = amplification(abi.MPD_MODEL_PP_D2Q9, RHO_V, [1])
    k, lam = out[0]
    assert abs(lam - math.exp(-cfg.pp_tau * cfg.rt * k * k * cfg.dt)) < 5e-5


def test_d2q9pp_stable_about_both_phases(oracle_mod):
    for rho in (RHO_V, RHO_L, RHO_FREE):
        assert max_amplification(abi.MPD_MODEL_PP_D2Q9, rho) < 1.0, rho


def test_d2v16_unstable_about_both_phases(oracle_mod):
    assert max_amplification(abi.MPD_MODEL_PP_D2V16, RHO_V) > 1.001
    assert max_amplification(abi.MPD_MODEL_PP_D2V16, RHO_L) > 1.01


def test_d2v16_unstable_without_the_force(oracle_mod):
    """Where dp/drho = RT the force has no linear part, so what grows is the
    16-velocity kinetic step itself."""
    h = 1e-7
    assert abs((_cs_eos(RHO_FREE + h, 0.8) - _cs_eos(RHO_FREE - h, 0.8)) / (2 * h) - 1 / 3) < 1e-6
    _, diag = amplification(abi.MPD_MODEL_PP_D2V16, RHO_FREE, range(1, 8), nx=16, diag=True)
    assert max(l for _, l in diag) > 1.02


@pytest.mark.gpu
def test_d2v16_liquid_vapour_512_diverges_in_both_precisions():
    """The liquid-vapour workload as specified -- D2V16, Carnahan-Starling at
    T/Tc = 0.8, 512^2 periodic, rho0 = 0.1645 (1 +- 1 %) per cell, FP32-bulk --
    beside the plain FP64 path from the same start: the two agree while the
    state is bounded (density within 1e-5 relative L2 after 50 steps) and both
    leave it (non-finite densities) well before the workload's 5e4 steps, as
    the linearised step predicts (amplification up to 1.03 per step about both
    phases).  The model, not the FP32 arithmetic, sets the horizon; the workload
    itself runs on D2Q9PP (tests/test_gpu_pp_observables.py)."""
    import numpy as np
    from mpdugks_amd import solver
    out = {}
    for prec in (abi.MPD_FP32_BULK, abi.MPD_FP64):
        cfg = default_config(abi.MPD_MODEL_PP_D2V16, nx=512, ny=512, pp_tr=0.8)
        cfg.precision = prec
        s = solver.Solver(cfg)
        s.init_pp_noise(0.1645, amp=0.01, seed=7)
        s.step(50)
        rho50 = s.download().m[abi.M["rho"]]
        first_bad = None
        for n in range(1, 9):
            s.step(250)
            r = s.reductions()
            if r["nonfinite"] > 0:
                first_bad = 50 + 250 * n
                break
        s.close()
        out[prec] = (rho50, first_bad)
    a, b = out[abi.MPD_FP32_BULK][0], out[abi.MPD_FP64][0]
    err = float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))
    print("D2V16 512^2 T/Tc 0.8: FP32 vs FP64 after 50 steps %.2e; non-finite by step FP32 %s, FP64 %s"
          % (err, out[abi.MPD_FP32_BULK][1], out[abi.MPD_FP64][1]))
    assert err < 1e-5
    assert out[abi.MPD_FP32_BULK][1] is not None and out[abi.MPD_FP64][1] is not None
