"""ctypes front end of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the reference results are compared
against or timed beside, never as the product.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-f", os.path.join(HERE, "Makefile")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        _lib.ora_create.restype = C.c_void_p
        _lib.ora_create.argtypes = [C.c_void_p]
        _lib.ora_destroy.argtypes = [C.c_void_p]
        for fn in (_lib.ora_set_state, _lib.ora_get_state):
            fn.argtypes = [C.c_void_p] + [C.c_void_p] * 5
            fn.restype = C.c_int
        _lib.ora_step.argtypes = [C.c_void_p, C.c_int]
        _lib.ora_step.restype = C.c_int
        _lib.ora_wall_flux_residual.argtypes = [C.c_void_p]
        _lib.ora_wall_flux_residual.restype = C.c_double
        _lib.ora_cell_phases.argtypes = [C.c_void_p]
        _lib.ora_cell_phases.restype = C.c_int
        _lib.ora_d2v16_tables.argtypes = [C.c_void_p] * 4
        _lib.ora_d2v16_tables.restype = None
        _lib.ora_d2v16_feq.argtypes = [C.c_double] * 4 + [C.c_void_p]
        _lib.ora_d2v16_feq.restype = None
        _lib.ora_set_face_order.argtypes = [C.c_void_p, C.c_int]
        _lib.ora_set_face_order.restype = C.c_int
        _lib.ora_set_one_copy.argtypes = [C.c_void_p, C.c_int]
        _lib.ora_set_one_copy.restype = C.c_int
        for fn in (_lib.ora_set_face_extra, _lib.ora_get_face_extra):
            fn.argtypes = [C.c_void_p] * 4
            fn.restype = C.c_int
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Oracle:
    """FP64 CPU restatement of the reference step on an nx*ny grid."""

    def __init__(self, cfg):
        from mpdugks_amd.abi import NM
        self.cfg = cfg
        self.N = cfg.nx * cfg.ny
        self.nm = NM
        self._h = lib().ora_create(C.byref(cfg))
        if not self._h:
            raise ValueError("oracle: unsupported configuration")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().ora_destroy(self._h)
            self._h = None

    def set_state(self, h, f, m, ffx=None, ffy=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64)
                for a in (h, f, m, ffx, ffy)]
        lib().ora_set_state(self._h, *[_ptr(a) for a in arrs])

    def get_state(self):
        from mpdugks_amd.abi import num_q
        N, Q = self.N, num_q(self.cfg.model)
        h = np.zeros((Q, N)); f = np.zeros((Q, N)); m = np.zeros((self.nm, N))
        ffx = np.zeros(2 * N); ffy = np.zeros(2 * N)
        lib().ora_get_state(self._h, _ptr(h), _ptr(f), _ptr(m), _ptr(ffx), _ptr(ffy))
        return h, f, m, ffx, ffy

    def step(self, n=1):
        lib().ora_step(self._h, int(n))

    def cell_phases(self):
        """Phases 9-11 (macros and sources) on the current populations."""
        lib().ora_cell_phases(self._h)

    def set_face_order(self, reference):
        """Pseudopotential D2Q9: True = visit faces in the reference's
        one-thread order (Ark/Cpp/DUGKS2DSolver.cpp:219-230), False = Jacobi."""
        if lib().ora_set_face_order(self._h, 1 if reference else 0) != 0:
            raise ValueError("reference face order: pseudopotential D2Q9, periodic only")

    def set_one_copy(self, one):
        """Test knob: Allen-Cahn UW with one face per location (without the
        second stored copy of each periodic face the reference keeps)."""
        lib().ora_set_one_copy(self._h, 1 if one else 0)

    def set_face_extra(self, psi_old=None, copy_fx=None, copy_fy=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64)
                for a in (psi_old, copy_fx, copy_fy)]
        lib().ora_set_face_extra(self._h, *[_ptr(a) for a in arrs])

    def get_face_extra(self):
        """(face Psi of the last step [2N], second-copy forces x [nx+ny], y [nx+ny])."""
        psi = np.zeros(2 * self.N)
        n = self.cfg.nx + self.cfg.ny
        cx = np.zeros(n); cy = np.zeros(n)
        lib().ora_get_face_extra(self._h, _ptr(psi), _ptr(cx), _ptr(cy))
        return psi, cx, cy

    def wall_flux_residual(self):
        return lib().ora_wall_flux_residual(self._h)


def d2v16_tables():
    """(A, xi_x, xi_y, E) of the oracle's D2V16 set: A[j][k] maps the 16
    equilibrium moments to populations, E_j = |xi_j|^2 (+ eta^2 on shell 0)."""
    a = np.zeros((16, 16)); x = np.zeros(16); y = np.zeros(16); e = np.zeros(16)
    lib().ora_d2v16_tables(_ptr(a), _ptr(x), _ptr(y), _ptr(e))
    return a, x, y, e


def d2v16_feq(rho, u, v, rt):
    """The oracle's D2V16 equilibrium populations at (rho, u, v), temperature rt."""
    f = np.zeros(16)
    lib().ora_d2v16_feq(rho, u, v, rt, _ptr(f))
    return f
