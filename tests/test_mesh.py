"""Fluent .cas meshes and the CellArray <-> structured-index map (mpdugks_amd/mesh.py).

The reference numbers cells in mesh-file order (CellArray) and writes its flow
field in that order; the fixture holds the SHA-256 of the reference's step-0
flow field for a mesh whose cells are numbered in a shuffled order (seed 7), so
the map read back from that file, handed to the library's writer, must give the
same bytes.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from mpdugks_amd import abi, mesh, solver
from mpdugks_amd.config import default_config

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("nx,ny,walls", [(16, 12, False), (24, 40, True)])
def test_round_trip_structured(tmp_path, nx, ny, walls):
    p = tmp_path / "m.cas"
    mesh.write_cas(str(p), nx, ny, walls_y=walls, lx=2.0 * nx, ly=2.0 * ny)
    m = mesh.read_cas(str(p))
    assert (m["nx"], m["ny"], m["dx"], m["dy"], m["walls_y"]) == (nx, ny, 2.0, 2.0, walls)
    assert np.array_equal(m["order"], np.arange(nx * ny))
    assert m["zones"][3] == mesh.BC_INTERIOR
    assert m["zones"][6] == (mesh.BC_WALL if walls else mesh.BC_PERIODIC)


def test_round_trip_shuffled(tmp_path):
    perm = np.random.default_rng(3).permutation(20 * 14)
    p = tmp_path / "m.cas"
    mesh.write_cas(str(p), 20, 14, cell_order=perm)
    assert np.array_equal(mesh.read_cas(str(p))["order"], perm)


def test_flowfield_in_reference_cell_order(tmp_path):
    with open(os.path.join(HERE, "golden", "ref_writers.json")) as fh:
        gold = json.load(fh)["permuted"]
    perm = np.random.default_rng(gold["seed"]).permutation(128 * 128)
    p = tmp_path / "m.cas"
    mesh.write_cas(str(p), 128, 128, cell_order=perm)
    order = mesh.read_cas(str(p))["order"]
    cfg = default_config(abi.MPD_MODEL_AC_HCZ)
    st = solver.ac_drop(cfg)
    out = tmp_path / "ff.dat"
    solver.write_flowfield(str(out), cfg, st, solver.output_info(cfg), order=order)
    with open(out, "rb") as fh:
        assert hashlib.sha256(fh.read()).hexdigest() == gold["sha256"]


def test_rejects_non_cartesian(tmp_path):
    p = tmp_path / "m.cas"
    mesh.write_cas(str(p), 8, 8)
    txt = p.read_text().replace("1.000000000000e+00 0.000000000000e+00", "1.250000000000e+00 0.000000000000e+00", 1)
    p.write_text(txt)
    with pytest.raises(ValueError):
        mesh.read_cas(str(p))
