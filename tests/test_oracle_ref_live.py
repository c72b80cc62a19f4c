"""Oracle against the compiled reference run live (oracle/_ref), for the parts
that need full reference fields as input.  Skipped where the reference sources
(and so the compiled reference) are absent.

* UW face scheme (reference built with the UW flag set), periodic and with
  bounce-back or NEE walls: bit-exact from the reference's own state over 10
  steps (and from the initialiser over 100, tests/golden/ref_*_uw*.json).  The
  reference's two stored copies of each periodic face (zones 12 and 8, each
  owned by its mesh cell) round apart, because for xi.n = 0 its UW average adds
  the owner first (Ark/Cpp/DUGKS2DSolver.cpp:580-603); the oracle carries the
  second copy's flux for the last column and row.
* Pseudopotential: the reference's face force reads neighbour-face Psi inside
  the loop that writes them (Ark/Cpp/DUGKS2DSolver.cpp:219-230, 733-746), so a
  whole step is order dependent.  The cell phases (Update_MacroVar with the
  Carnahan-Starling EOS and Psi, then Update_PseudoForce(cell)) are pinned bit
  for bit by feeding the oracle the reference's populations; a whole step from
  the reference's own state after one step (face Psi and forces included) is
  bit-exact in the oracle's reference face order, and in Jacobi order lands
  1.1e-4 away.  The face numbering and stencils the reference-order mode
  assumes are checked against the reference's own face topology.
"""
import os

import numpy as np
import pytest

from mpdugks_amd import abi
from mpdugks_amd.config import default_config

ref_run = pytest.importorskip("ref_run")
pytestmark = pytest.mark.skipif(not ref_run.available(), reason="reference sources not present")

KEYS = {"phi": "Phi", "rho": "Rho", "u": "U", "v": "V", "p": "p", "phix": "Phi_x", "phiy": "Phi_y",
        "rhox": "Rho_x", "rhoy": "Rho_y", "fx": "Fx", "fy": "Fy", "lap": "laplacianPhi", "ppu": "prevPhiU",
        "ppv": "prevPhiV", "psi": "Psi", "scg": "scG", "mu": "Mu", "tau": "ftau"}
N = 128 * 128


def to_state(d):
    m = np.zeros((abi.NM, N))
    for a, b in KEYS.items():
        m[abi.M[a]] = d[b].reshape(N)
    h = d["h"].reshape(9, N) if "h" in d else np.zeros((9, N))
    return h, d["f"].reshape(9, N), m


def face_forces(path):
    raw = np.fromfile(path, dtype=np.float64, offset=8).reshape(-1, 12)
    fx, fy = np.zeros(2 * N), np.zeros(2 * N)
    for xf, yf, vx, vy, _bc, _r, _u, _v, _p, _psi, Fx, Fy in raw:
        if abs(vy) < 0.5:
            idx = int(round(yf - 0.5)) * 128 + int(round(xf)) % 128
        else:
            idx = N + (int(round(yf)) % 128) * 128 + int(round(xf - 0.5))
        fx[idx], fy[idx] = Fx, Fy
    return fx, fy


UW_LIVE = [("hcz_uw", abi.MPD_MODEL_AC_HCZ, abi.MPD_BC_PERIODIC),
           ("acp_uw", abi.MPD_MODEL_AC_P, abi.MPD_BC_PERIODIC),
           ("hcz_uw_wall", abi.MPD_MODEL_AC_HCZ, abi.MPD_BC_WALL_BB),
           ("acp_uw_nee", abi.MPD_MODEL_AC_P, abi.MPD_BC_WALL_NEE)]


@pytest.mark.parametrize("name,model,bc", UW_LIVE, ids=[c[0] for c in UW_LIVE])
def test_uw_against_reference(name, model, bc, oracle_mod):
    out = ref_run.run(name, [0, 1, 10], 10)
    cfg = default_config(model)
    cfg.flux, cfg.bc_y = abi.MPD_FLUX_UW, bc
    o = oracle_mod.Oracle(cfg)
    o.set_state(*to_state(ref_run.load_cells(os.path.join(out, "cells_step0.bin"))))
    done = 0
    for k in (1, 10):
        o.step(k - done)
        done = k
        h, f, m, _, _ = o.get_state()
        rh, rf, rm = to_state(ref_run.load_cells(os.path.join(out, "cells_step%d.bin" % k)))
        assert np.array_equal(h, rh) and np.array_equal(f, rf) and np.array_equal(m, rm), k


def test_uw_second_copy_matters(oracle_mod):
    """The second copy is not redundant: from the reference's step-1 state, one
    periodic UW step with one face per location (the oracle's test knob; the
    last column and row then take the first copy's flux) misses the
    reference's step in the last bits, and the default oracle hits it."""
    out = ref_run.run("hcz_uw", [1, 2], 2)
    cfg = default_config(abi.MPD_MODEL_AC_HCZ)
    cfg.flux = abi.MPD_FLUX_UW
    s1 = to_state(ref_run.load_cells(os.path.join(out, "cells_step1.bin")))
    r2 = to_state(ref_run.load_cells(os.path.join(out, "cells_step2.bin")))
    got = {}
    for one in (False, True):
        o = oracle_mod.Oracle(cfg)
        o.set_state(*s1)
        o.set_one_copy(one)
        o.step(1)
        got[one] = o.get_state()[:3]
    assert all(np.array_equal(a, b) for a, b in zip(got[False], r2))
    assert not all(np.array_equal(a, b) for a, b in zip(got[True], r2))
    dev = max(float(np.abs(a - b).max() / np.abs(b).max()) for a, b in zip(got[True], r2))
    assert dev < 1e-14, dev


def test_pseudopotential_cell_phases_bit_exact(oracle_mod):
    out = ref_run.run("pp", [0, 1, 2], 2)
    cfg = default_config(abi.MPD_MODEL_PP_D2Q9)
    for k in (0, 1):
        _, _, mk = to_state(ref_run.load_cells(os.path.join(out, "cells_step%d.bin" % k)))
        _, fk1, mk1 = to_state(ref_run.load_cells(os.path.join(out, "cells_step%d.bin" % (k + 1))))
        o = oracle_mod.Oracle(cfg)
        o.set_state(np.zeros((9, N)), fk1, mk)
        o.cell_phases()
        _, _, m, _, _ = o.get_state()
        for name in ("rho", "u", "v", "p", "psi", "scg", "fx", "fy"):
            assert np.array_equal(m[abi.M[name]], mk1[abi.M[name]]), (k, name)


def face_psi(path):
    raw = np.fromfile(path, dtype=np.float64, offset=8).reshape(-1, 12)
    psi = np.zeros(2 * N)
    for xf, yf, vx, vy, _bc, _r, _u, _v, _p, ps, _fx, _fy in raw:
        if abs(vy) < 0.5:
            idx = int(round(yf - 0.5)) * 128 + int(round(xf)) % 128
        else:
            idx = N + (int(round(yf)) % 128) * 128 + int(round(xf - 0.5))
        psi[idx] = ps
    return psi


def test_pseudopotential_step_from_reference_state(oracle_mod):
    out = ref_run.run("pp", [1, 2], 2)
    cfg = default_config(abi.MPD_MODEL_PP_D2Q9)
    _, f1, m1 = to_state(ref_run.load_cells(os.path.join(out, "cells_step1.bin")))
    _, f2, m2 = to_state(ref_run.load_cells(os.path.join(out, "cells_step2.bin")))
    ffx, ffy = face_forces(os.path.join(out, "faces_step1.bin"))
    psi1 = face_psi(os.path.join(out, "faces_step1.bin"))
    res = {}
    for ref in (True, False):
        o = oracle_mod.Oracle(cfg)
        o.set_state(np.zeros((9, N)), f1, m1, ffx, ffy)
        o.set_face_extra(psi_old=psi1)
        o.set_face_order(ref)
        o.step(1)
        res[ref] = o.get_state()
    _, f, m, _, _ = res[True]
    assert np.array_equal(f, f2) and np.array_equal(m, m2)
    _, f, m, _, _ = res[False]
    rel = np.abs(f - f2).max() / np.abs(f2).max()
    assert 9e-5 < rel < 1.4e-4, rel   # the Jacobi order's stale-neighbour effect
    assert abs(m[abi.M["rho"]].sum() - m2[abi.M["rho"]].sum()) < 1e-12 * m2[abi.M["rho"]].sum()


def test_face_numbering_matches_reference_topology(oracle_mod):
    """The reference-order mode's face ranks and stencil neighbours (oracle
    face_rank / face_obj_step) against the reference's own FaceArray, faceFaces
    and visiting lists (Ark/Cpp/Mesh/MeshCartesian.cpp:296-368)."""
    os.environ["MPD_REF_TOPO"] = "1"
    try:
        out = ref_run.run("pp", [0], 0)
    finally:
        os.environ.pop("MPD_REF_TOPO")
    raw = open(os.path.join(out, "topo.bin"), "rb").read()
    nf, ni, npf, nw = np.frombuffer(raw[:16], np.int32)
    rec = np.dtype([("ff", np.int32, 9), ("shadow", np.int32), ("bc", np.int32), ("d", np.float64, 6)])
    faces = np.frombuffer(raw[16:16 + nf * rec.itemsize], rec)
    order = np.frombuffer(raw[16 + nf * rec.itemsize:], np.int32)
    assert np.array_equal(order, np.arange(ni + npf)) and nw == 0
    nx = ny = 128

    def obj(k):  # face object (dir, i, j) of FaceArray index k, from its centre
        xf, yf = faces[k]["d"][0], faces[k]["d"][1]
        if abs(xf - round(xf)) < 1e-9:
            return 0, int(round(xf)), int(yf)
        return 1, int(xf), int(round(yf))

    def rank(d, i, j):  # oracle face_rank
        NX, NI = ny * (nx - 1), ny * (nx - 1) + (ny - 1) * nx
        if d == 0:
            return NI + j if i == 0 else (NI + ny + j if i == nx else j * (nx - 1) + i - 1)
        return NI + 2 * ny + i if j == 0 else (NI + 2 * ny + nx + i if j == ny else NX + (j - 1) * nx + i)

    def step(d, i, j, di, dj):  # oracle face_obj_step
        a, b = i + di, j + dj
        if d == 0:
            a = a + nx if a < 0 else (a - nx if a > nx else a)
            b %= ny
        else:
            a %= nx
            b = b + ny if b < 0 else (b - ny if b > ny else b)
        return a, b
    offs = {1: (1, 0), 2: (1, 1), 3: (0, 1), 4: (-1, 1), 5: (-1, 0), 6: (-1, -1), 7: (0, -1), 8: (1, -1)}
    for k in range(nf):
        d, i, j = obj(k)
        assert rank(d, i, j) == k
        for s, (di, dj) in offs.items():
            assert rank(d, *step(d, i, j, di, dj)) == faces[k]["ff"][s], (k, s)


def test_integration_snippet_compiles_against_reference_headers(tmp_path):
    """The driver-side replacement of DUGKS2DSolver() shown in INTEGRATION.md
    compiles (g++ -fsyntax-only) against the reference's own headers and
    include/mpdugks.h.  The prelude declares what the reference's
    DUGKS2DSolver.cpp itself declares at file scope."""
    import re
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    text = open(os.path.join(root, "INTEGRATION.md")).read()
    block = next(b for b in re.findall(r"```cpp\n(.*?)```", text, re.S) if "void DUGKS2DSolver()" in b)
    prelude = ("#include \"DUGKSDeclaration.h\"\n"
               "static double ResidualPer1k = 1.0;\n"
               "void Output_Flowfield(double const &t, int step);\n"
               "void Output_Diag(int step);\nvoid Output_MidX(int step);\n")
    src = tmp_path / "driver.cpp"
    src.write_text(prelude + block)
    ark = ref_run.REF_ROOT
    subprocess.run(["g++", "-std=c++11", "-fsyntax-only", "-w", "-I" + os.path.join(ark, "Header"),
                    "-I" + os.path.join(ark, "Lib"), "-I" + os.path.join(ark, "Constant"),
                    "-I" + os.path.join(root, "include"), str(src)], check=True)


def test_reference_pseudopotential_with_walls_cannot_run():
    """The library refuses pseudopotential handles with walls (mpd_create:
    MPD_ERR_INVALID), as the reference cannot run them: with walls on,
    CreatMesh.h:11-13 skips CarfaceFacesConstruct, so every face's faceFaces
    stays null and the face force (GradSchemeBasic.h:181-208) dereferences it.
    The reference built with the pseudopotential flags and bounce-back walls
    dies with SIGSEGV in its first step."""
    import subprocess
    with pytest.raises(subprocess.CalledProcessError) as e:
        ref_run.run("pp_wall", [1], 1, timeout=120)
    assert e.value.returncode == -11
    cfg = default_config(abi.MPD_MODEL_PP_D2Q9, nx=32, ny=32)
    cfg.bc_y = abi.MPD_BC_WALL_BB
    import ctypes as C
    from mpdugks_amd import solver
    assert solver.lib().mpd_create(C.byref(cfg), C.byref(C.c_void_p())) == -1
