"""C ABI checks that need no GPU: the library loads, exports every function
include/mpdugks.h declares, and its host-only entry points agree with the
Python mirror (defaults) and with the reference (initial state)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from mpdugks_amd import abi, solver
from mpdugks_amd.config import default_config

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "mpdugks.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*|void)\s*\*?\s*(mpd_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = solver.lib()
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert missing == []


def test_python_mirror_binds_every_symbol():
    assert set(declared_functions()) <= set(solver._SYMBOLS)


def test_abi_version():
    assert solver.lib().mpd_abi_version() == 1


def test_default_config_matches_python_mirror():
    for m in (abi.MPD_MODEL_AC_P, abi.MPD_MODEL_AC_HCZ, abi.MPD_MODEL_PP_D2Q9):
        assert bytes(solver.default_config(m)) == bytes(default_config(m))


def test_config_struct_layout():
    # 6 int32 then 19 doubles, no padding surprises
    assert C.sizeof(abi.MpdConfig) == 6 * 4 + 19 * 8
    assert abi.MpdConfig.dx.offset == 24


def test_invalid_init_arguments_are_rejected():
    cfg = default_config(abi.MPD_MODEL_PP_D2Q9)
    rc = solver.lib().mpd_init_ac_drop(C.byref(cfg), 1.0, 1.0, 1.0, 1, 1.0, 0.1, None, None, None)
    assert rc == -1


def test_interface_initialiser_profile():
    cfg = default_config(abi.MPD_MODEL_AC_HCZ, nx=8, ny=40)
    y0 = np.full(8, 20.0)
    st = solver.ac_interface(cfg, y0, sign=+1)
    phi = st.field("phi")
    assert phi[0, 0] < 1e-3 and phi[-1, 0] > 1 - 1e-3           # phase L above
    assert np.allclose(st.h.sum(axis=0), st.m[abi.M["phi"]], rtol=0, atol=1e-15)


def test_fp32_bulk_configurations_pass_the_checks():
    """FP32-bulk exists for the Allen-Cahn models (CD and UW faces; periodic,
    bounce-back or NEE y), the pseudopotential D2Q9 model (CD and UW, periodic)
    and D2V16 (CD, periodic): every such configuration passes the checks (on a
    machine without a GPU mpd_create then fails with MPD_ERR_DEVICE, -2), and
    D2V16 with UW faces is refused with MPD_ERR_INVALID before any device call."""
    L = solver.lib()
    ok = [(abi.MPD_MODEL_PP_D2V16, abi.MPD_FLUX_CD), (abi.MPD_MODEL_PP_D2Q9, abi.MPD_FLUX_CD),
          (abi.MPD_MODEL_PP_D2Q9, abi.MPD_FLUX_UW)]
    for model, flux in ok:
        cfg = default_config(model, nx=32, ny=32)
        cfg.precision, cfg.flux = abi.MPD_FP32_BULK, flux
        h = C.c_void_p()
        rc = L.mpd_create(C.byref(cfg), C.byref(h))
        if rc == 0:
            L.mpd_destroy(h)
        assert rc in (0, -2), (model, flux, rc, L.mpd_last_error())
    cfg = default_config(abi.MPD_MODEL_PP_D2V16, nx=32, ny=32)
    cfg.precision, cfg.flux = abi.MPD_FP32_BULK, abi.MPD_FLUX_UW
    h = C.c_void_p()
    assert L.mpd_create(C.byref(cfg), C.byref(h)) == -1
    assert b"D2V16" in L.mpd_last_error()


def test_header_compiles_as_c99_and_links(tmp_path):
    """include/mpdugks.h as a C99 header: a C translation unit compiled with
    -std=c99 -pedantic -Werror, linked against libmpdugks_hip.so, runs the
    host-only entry points through it."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = os.path.join(root, "tests", "c", "abi_smoke.c")
    libdir = os.path.dirname(solver.LIB_PATH)
    exe = str(tmp_path / "abi_smoke")
    subprocess.run(["gcc", "-std=c99", "-pedantic", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(root, "include"),
                    src, "-o", exe, "-L" + libdir, "-lmpdugks_hip", "-Wl,-rpath," + libdir], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    mass = float(out.split()[-1])
    cfg = default_config(abi.MPD_MODEL_AC_HCZ, nx=32, ny=24)
    st = solver.ac_drop(cfg, cx=16.0, cy=12.0, radius=8.0, sign=-1)
    assert abs(mass - st.m[abi.M["phi"]].sum()) < 1e-6


def test_fp32_mesh_beyond_buffer_range_refused():
    """FP32-bulk Allen-Cahn handles address a state set through 32-bit buffer
    resources: a mesh whose float planes would span 4 GiB or more is refused
    with MPD_ERR_INVALID by mpd_create, before any device call."""
    # HCZ: 22 planes, at most 48,806,446 cells; 8192 x 5958 is the first row
    # over (8192 x 5957 runs: tests/test_gpu_max_mesh.py)
    for model, nx, ny in ((abi.MPD_MODEL_AC_HCZ, 7000, 7000), (abi.MPD_MODEL_AC_P, 8000, 8000),
                          (abi.MPD_MODEL_AC_HCZ, 8192, 5958)):
        cfg = default_config(model, nx=nx, ny=ny)
        cfg.precision = abi.MPD_FP32_BULK
        h = C.c_void_p()
        rc = solver.lib().mpd_create(C.byref(cfg), C.byref(h))
        assert rc == -1 and not h.value
        assert b"too large for the FP32-bulk" in solver.lib().mpd_last_error()
    # pseudopotential: a set's float block is q + 4 planes (D2Q9 13, D2V16 20)
    for model, nx, ny in ((abi.MPD_MODEL_PP_D2V16, 8192, 6554), (abi.MPD_MODEL_PP_D2Q9, 9000, 9200)):
        cfg = default_config(model, nx=nx, ny=ny)
        cfg.precision = abi.MPD_FP32_BULK
        h = C.c_void_p()
        rc = solver.lib().mpd_create(C.byref(cfg), C.byref(h))
        assert rc == -1 and not h.value
        assert b"too large for the FP32-bulk pseudopotential" in solver.lib().mpd_last_error()


def test_fp32_uw_configurations():
    """FP32-bulk UW faces: the Allen-Cahn models with periodic, bounce-back and
    NEE y and the periodic pseudopotential D2Q9 model pass the checks (on a
    machine without a GPU mpd_create then fails with MPD_ERR_DEVICE); D2V16 is
    refused with MPD_ERR_INVALID before any device call."""
    for model in (abi.MPD_MODEL_AC_HCZ, abi.MPD_MODEL_AC_P):
        for bc in (abi.MPD_BC_PERIODIC, abi.MPD_BC_WALL_BB, abi.MPD_BC_WALL_NEE):
            cfg = default_config(model, nx=48, ny=32)
            cfg.precision, cfg.flux, cfg.bc_y = abi.MPD_FP32_BULK, abi.MPD_FLUX_UW, bc
            h = C.c_void_p()
            rc = solver.lib().mpd_create(C.byref(cfg), C.byref(h))
            if rc == 0:
                solver.lib().mpd_destroy(h)
            assert rc in (0, -2), (model, bc, rc)
    cfg = default_config(abi.MPD_MODEL_PP_D2Q9, nx=48, ny=32)
    cfg.precision, cfg.flux = abi.MPD_FP32_BULK, abi.MPD_FLUX_UW
    h = C.c_void_p()
    rc = solver.lib().mpd_create(C.byref(cfg), C.byref(h))
    if rc == 0:
        solver.lib().mpd_destroy(h)
    assert rc in (0, -2), rc
    cfg = default_config(abi.MPD_MODEL_PP_D2V16, nx=48, ny=32)
    cfg.precision, cfg.flux = abi.MPD_FP32_BULK, abi.MPD_FLUX_UW
    assert solver.lib().mpd_create(C.byref(cfg), C.byref(C.c_void_p())) == -1


def test_upload_checks_array_sizes_before_the_library():
    """A host state built for another mesh is refused on the host: the library
    would read nx*ny-sized planes past the end of the smaller arrays."""
    cfg = default_config(abi.MPD_MODEL_AC_HCZ, nx=16, ny=12)
    s = solver.Solver.__new__(solver.Solver)   # no device handle: the check comes first
    s.cfg, s._h = cfg, C.c_void_p()
    with pytest.raises(solver.MpdError, match="mesh needs"):
        s.upload(solver.State(default_config(abi.MPD_MODEL_AC_HCZ, nx=8, ny=12)))
    pp = default_config(abi.MPD_MODEL_PP_D2Q9, nx=16, ny=12)
    s.cfg = pp
    st = solver.State(pp)
    st.ffx = st.ffx[:10]
    with pytest.raises(solver.MpdError, match="ffx holds 10 values"):
        s.upload(st)


def test_host_entry_points_report_why_they_fail():
    """The host initialisers and the writers set mpd_last_error as the device
    entry points do."""
    L = solver.lib()
    pp = default_config(abi.MPD_MODEL_PP_D2Q9, nx=8, ny=8)
    h = np.zeros((9, 64))
    rc = L.mpd_init_ac_drop(C.byref(pp), 4.0, 4.0, 2.0, -1, 1.0, 0.1, h.ctypes.data_as(C.c_void_p),
                            h.ctypes.data_as(C.c_void_p), np.zeros((abi.NM, 64)).ctypes.data_as(C.c_void_p))
    assert rc == -1
    assert b"mpd_init_ac_drop: starts the Allen-Cahn models" in L.mpd_last_error()
    info = solver.output_info(default_config(abi.MPD_MODEL_AC_HCZ, nx=8, ny=8))
    buf = C.create_string_buffer(8)
    assert L.mpd_output_name(abi.MPD_OUT_SUMRHO, C.byref(info), 0, 0.0, buf, 8) == -1
    assert b"the buffer 8" in L.mpd_last_error()
    rc = L.mpd_append_convergence(b"/nonexistent-dir/x.dat", 0.0, 1.0)
    assert rc == -1 and b"cannot open /nonexistent-dir/x.dat" in L.mpd_last_error()
