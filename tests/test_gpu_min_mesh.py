"""The smallest meshes the library accepts (4 cells a side) and other meshes
narrower than every tile's halo, on every path, against the oracle.

On a 4-cell periodic mesh each tile's halo (3 cells for the Allen-Cahn UW
tile, 4 for the pseudopotential UW face-Psi region) wraps around the mesh more
than once, and a 4-row walled mesh holds both walls' ghost rows in one tile:
the kernels take their general wrap there.  The FP64 paths (one kernel per
phase for every model, face scheme and boundary; the fused Allen-Cahn CD
kernel) equal the oracle bit for bit.  The FP32-bulk paths meet their bar
after 100 steps: rho, phi, p, Psi and the populations within 1e-5 relative L2
of the FP64 oracle, the velocity field (u, v) within 1e-3 of
max(|(u, v)|_L2, 1e-3 c_s sqrt(N)) with c_s = sqrt(RT), and the conserved sum
(phi, or rho for the pseudopotential) within 1e-11 of the oracle's.  Over that horizon FP64 runs
started from states perturbed at FP32 round-off (6e-8 relative, three seeds)
stay within 1e-8 of the unperturbed run in rho, phi, p and Psi on every one of
these cases.  The velocity is measured as one field, not per component: on
these meshes a drop spans the short side, so one component nearly vanishes by
symmetry (|v| 1.6e-4 against |u| 6.3e-2 for the pseudopotential drop on 9x4)
and carries only the rounding of a cancellation.  Its floor is FP32's
resolution of a velocity: a drop at rest carries only spurious currents (2.6e-9
in the FP64 run of the pseudopotential UW drop on 4x4), far below an FP32 ulp
of c_s (6e-8), so the bar is about ten such ulps per cell.
"""
import numpy as np
import pytest

from mpdugks_amd import abi, solver
from mpdugks_amd.config import default_config
from helpers import make_case, max_rel, oracle_run

pytestmark = pytest.mark.gpu

MESHES = [(4, 4), (4, 7), (9, 4)]
FLUX = {"cd": abi.MPD_FLUX_CD, "uw": abi.MPD_FLUX_UW}
AC = [(m, fl, bc) for m in ("ac_hcz", "ac_p") for fl in ("cd", "uw") for bc in ("periodic", "bb", "nee")]
PP = [("pp_d2q9", "cd", "periodic"), ("pp_d2q9", "uw", "periodic")]


def _case(model, fl, bc, nx, ny, precision=abi.MPD_FP64):
    if model == "pp_d2v16":
        cfg = default_config(abi.MPD_MODEL_PP_D2V16, nx=nx, ny=ny)
        return cfg, solver.scmp_drop(cfg, radius=0.4 * min(nx, ny))
    return make_case(model, nx=nx, ny=ny, flux=FLUX[fl], walls=bc != "periodic", nee=bc == "nee",
                     radius=0.4 * min(nx, ny), precision=precision)


def _ids(cases):
    return ["%s_%s_%s_%dx%d" % (m, fl, bc, nx, ny) for (m, fl, bc), (nx, ny) in cases]


def _grid(cases):
    return [(c, mesh) for c in cases for mesh in MESHES]


def _assert_equal(gpu, ref, tag):
    names = ["h", "f", "m"]
    if gpu.cfg.model == abi.MPD_MODEL_PP_D2Q9:
        names += ["ffx", "ffy"]  # the face forces are state only for this model
    for name in names:
        a, b = getattr(gpu, name), getattr(ref, name)
        if a is None:
            continue
        if not np.array_equal(a, b):
            bad = np.argwhere(a != b)
            raise AssertionError("%s: plane %s differs at %d entries (first %s), max rel %.3e"
                                 % (tag, name, len(bad), bad[0].tolist(), max_rel(a, b)))


def _run_exact(cfg, st, path, tag):
    steps = [1, 3, 20]
    ref = oracle_run(cfg, st, steps)
    s = solver.Solver(cfg, path=path)
    s.upload(st)
    done = 0
    for k in steps:
        s.step(k - done)
        done = k
        _assert_equal(s.download(), ref[k], "%s step %d" % (tag, k))
    s.close()


FP64_PHASES = _grid(AC + PP + [("pp_d2v16", "cd", "periodic")])


@pytest.mark.parametrize("case,mesh", FP64_PHASES, ids=_ids(FP64_PHASES))
def test_fp64_phases_bit_exact(case, mesh, oracle_mod):
    cfg, st = _case(*case, *mesh)
    _run_exact(cfg, st, abi.MPD_PATH_PHASES, "phases " + _ids([(case, mesh)])[0])


FP64_FUSED = _grid([c for c in AC if c[1] == "cd"])


@pytest.mark.parametrize("case,mesh", FP64_FUSED, ids=_ids(FP64_FUSED))
def test_fp64_fused_bit_exact(case, mesh, oracle_mod):
    cfg, st = _case(*case, *mesh)
    _run_exact(cfg, st, abi.MPD_PATH_FUSED, "fused " + _ids([(case, mesh)])[0])


def rel_l2(a, b, floor=0.0):
    return float(np.sqrt(((a - b) ** 2).sum()) / max(np.sqrt((b ** 2).sum()), floor))


FP32 = [(c, mesh, abi.MPD_PATH_FUSED) for c, mesh in _grid(AC + PP)]
FP32 += [(c, mesh, abi.MPD_PATH_PHASES) for c, mesh in _grid(PP[:1])]


@pytest.mark.parametrize("case,mesh,path", FP32,
                         ids=["%s_%s" % (i, "tile" if p == abi.MPD_PATH_FUSED else "phases")
                              for i, (_, _, p) in zip(_ids([(c, m) for c, m, _ in FP32]), FP32)])
def test_fp32_bulk_vs_fp64_oracle(case, mesh, path, oracle_mod):
    steps = 100
    cfg, st = _case(*case, *mesh, precision=abi.MPD_FP32_BULK)
    ref = oracle_run(cfg, st, [steps])[steps]
    s = solver.Solver(cfg, path=path)
    s.upload(st)
    s.step(steps)
    g = s.download()
    r = s.reductions()
    s.close()
    pp = case[0] == "pp_d2q9"
    n = mesh[0] * mesh[1]
    keys = ("rho", "psi") if pp else ("rho", "phi", "p")
    errs = {k: rel_l2(g.m[abi.M[k]], ref.m[abi.M[k]]) for k in keys}
    errs["f"] = rel_l2(g.f, ref.f)
    uv = lambda state: np.stack([state.m[abi.M["u"]], state.m[abi.M["v"]]])
    errs["uv"] = rel_l2(uv(g), uv(ref), floor=1e-3 * np.sqrt(cfg.rt * n))
    errs["uv_rms"] = float(np.sqrt(((uv(g) - uv(ref)) ** 2).sum() / n))
    cons = "rho" if pp else "phi"
    drift = abs(g.m[abi.M[cons]].sum() / ref.m[abi.M[cons]].sum() - 1)
    print("fp32 %s %s: %s, sum(%s) %.1e" % (_ids([(case, mesh)])[0], "tile" if path == abi.MPD_PATH_FUSED
                                            else "phases", errs, cons, drift))
    assert r["status"] == 0 and r["nonfinite"] == 0
    assert all(errs[k] < 1e-5 for k in keys + ("f",)), errs
    assert errs["uv"] < 1e-3, errs
    assert drift < 1e-11
