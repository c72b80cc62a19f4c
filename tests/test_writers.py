"""Output files of the reference driver (Ark/Cpp/Output.cpp), byte for byte.

Fixture tests/golden/ref_writers.json holds the names and SHA-256 of the files
the compiled reference wrote for the shipped AC+HCZ drop case (Case.ark; flow
field, MidX, MidY and Diag at steps 0 and 100) and the text of its SumRho and
Residual records at step 1000.  The states fed to the library's writers come
from the bit-exact initialiser and the bit-exact oracle (tests/test_oracle_golden.py
pins both), so every file must match exactly.  Host code only: no GPU.
"""
import ctypes as C
import hashlib
import json
import os

import numpy as np
import pytest

from mpdugks_amd import abi, solver
from mpdugks_amd.config import default_config

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "ref_writers.json")) as fh:
    GOLD = json.load(fh)


def _sha(path):
    with open(path, "rb") as fh:
        return hashlib.sha256(fh.read()).hexdigest()


def _case():
    cfg = default_config(abi.MPD_MODEL_AC_HCZ, nx=128, ny=128)
    return cfg, solver.ac_drop(cfg)


@pytest.fixture(scope="module")
def states():
    from helpers import oracle_run
    cfg, st = _case()
    out = oracle_run(cfg, st, [100, 1000])
    return cfg, {0: st, 100: out[100], 1000: out[1000]}


def test_case_file(tmp_path):
    cfg, _ = _case()
    p = tmp_path / "Case.ark"
    solver.write_case(str(p), solver.case_info(cfg))
    assert _sha(p) == GOLD["files"]["Case.ark"]["sha256"], p.read_text()


@pytest.mark.parametrize("step", [0, 100])
def test_flowfield_and_profiles(tmp_path, states, step):
    cfg, st = states
    info = solver.output_info(cfg)
    kinds = [abi.MPD_OUT_FLOWFIELD, abi.MPD_OUT_MIDX, abi.MPD_OUT_MIDY, abi.MPD_OUT_DIAG]
    for kind in kinds:
        name = solver.output_name(kind, info, step, step * cfg.dt)
        assert name in GOLD["files"], name
        p = tmp_path / os.path.basename(name)
        if kind == abi.MPD_OUT_FLOWFIELD:
            solver.write_flowfield(str(p), cfg, st[step], info)
        else:
            solver.write_profile(str(p), kind, cfg, st[step], info)
        assert os.path.getsize(p) == GOLD["files"][name]["bytes"], name
        assert _sha(p) == GOLD["files"][name]["sha256"], name


def test_convergence_records(tmp_path, states):
    """Update_SumRho / Update_Residual at step 1000 (DUGKS2DSolver.cpp:793-846):
    SumRho = sum(rho) - rho of CarCellArray[64][65], summed in cell order; the first
    residual has Rho_1k = 0."""
    cfg, st = states
    info = solver.output_info(cfg)
    rho = st[1000].m[abi.M["rho"]]
    s = 0.0
    for v in rho:
        s += float(v)
    s -= float(rho[64 * cfg.nx + 63])
    sq = 0.0
    for v in rho:
        sq += float(v) * float(v)
    res = float(np.sqrt(sq / (sq + 1.0e-30)))
    for kind, value in ((abi.MPD_OUT_SUMRHO, s), (abi.MPD_OUT_RESIDUAL, res)):
        name = solver.output_name(kind, info, 1000, 1000 * cfg.dt)
        p = tmp_path / os.path.basename(name)
        solver.append_convergence(str(p), 1000 * cfg.dt, value)
        assert p.read_text() == GOLD["text"][name], name


def test_flowfield_cell_order(tmp_path, states):
    """A driver whose CellArray is not in structured order passes the map."""
    cfg, st = states
    info = solver.output_info(cfg)
    order = np.random.default_rng(1).permutation(cfg.nx * cfg.ny)
    p = tmp_path / "perm.dat"
    solver.write_flowfield(str(p), cfg, st[0], info, order=order)
    lines = p.read_text().split("\n")
    rho = st[0].m[abi.M["rho"]]
    assert lines[0] == "(300 (101 2 1 0 0 1 16384)("
    assert float(lines[1]) == float("%.12e" % rho[order[0]])
    assert float(lines[5]) == float("%.12e" % rho[order[4]])


def test_unterminated_text_fields_are_read_to_their_end():
    """A text field that fills its array has no NUL: the writers read it to the
    end of the array and no further (here mesh_file, 128 letters, and mesh_type,
    32, followed in the struct by nothing that may be read as text)."""
    cfg = default_config(abi.MPD_MODEL_AC_HCZ, nx=32, ny=32)
    info = solver.output_info(cfg)
    C.memmove(C.addressof(info) + type(info).mesh_file.offset, b"m" * 128, 128)
    C.memmove(C.addressof(info) + type(info).mesh_type.offset, b"t" * 32, 32)
    name = solver.output_name(abi.MPD_OUT_SUMRHO, info, 0, 0.0)
    assert name.endswith("_" + "m" * 128 + ".dat")
    name = solver.output_name(abi.MPD_OUT_FLOWFIELD, info, 5, 1.0)
    assert "_" + "t" * 32 + str(info.mesh_nl) + "_CFL" in name
