"""Python host front end of the step library, mirroring the reference's interface.

The reference exposes one entry point for the hot path, ``DUGKS2DSolver()``
(Ark/Cpp/DUGKS2DSolver.cpp:148), run after an initialiser such as
``AC_Drop()`` (Ark/Cpp/Preprocess.cpp:705).  ``Solver`` keeps those names
(``ac_drop``, ``scmp_drop``, ``step``) over the C ABI of include/mpdugks.h.
Every call goes to libmpdugks_hip.so; there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from . import abi
from .abi import MpdConfig, MpdSums, NM, M

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libmpdugks_hip.so")
_lib = None

_SYMBOLS = {
    "mpd_abi_version": ([], C.c_int),
    "mpd_default_config": ([C.c_int32, C.c_void_p], C.c_int),
    "mpd_create": ([C.c_void_p, C.c_void_p], C.c_int),
    "mpd_destroy": ([C.c_void_p], C.c_int),
    "mpd_set_path": ([C.c_void_p, C.c_int32], C.c_int),
    "mpd_set_tile": ([C.c_void_p, C.c_int32], C.c_int),
    "mpd_set_face_order": ([C.c_void_p, C.c_int32], C.c_int),
    "mpd_upload_face_psi": ([C.c_void_p, C.c_void_p], C.c_int),
    "mpd_download_face_psi": ([C.c_void_p, C.c_void_p], C.c_int),
    "mpd_get_tile": ([C.c_void_p, C.POINTER(C.c_int32)], C.c_int),
    "mpd_upload_state": ([C.c_void_p] + [C.c_void_p] * 5, C.c_int),
    "mpd_download_state": ([C.c_void_p] + [C.c_void_p] * 5, C.c_int),
    "mpd_step": ([C.c_void_p, C.c_int64], C.c_int),
    "mpd_reductions": ([C.c_void_p, C.c_int, C.c_void_p], C.c_int),
    "mpd_synchronize": ([C.c_void_p], C.c_int),
    "mpd_device_bytes": ([C.c_void_p, C.c_void_p], C.c_int),
    "mpd_kernel_timing": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "mpd_last_error": ([], C.c_char_p),
    "mpd_set_kernel_timing": ([C.c_void_p, C.c_int], C.c_int),
    "mpd_batch_create": ([C.c_void_p], C.c_int),
    "mpd_batch_destroy": ([C.c_void_p], C.c_int),
    "mpd_batch_add": ([C.c_void_p, C.c_void_p], C.c_int),
    "mpd_batch_remove": ([C.c_void_p, C.c_void_p], C.c_int),
    "mpd_batch_size": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "mpd_batch_step": ([C.c_void_p, C.c_int64], C.c_int),
    "mpd_batch_timing": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "mpd_batch_set_tile": ([C.c_void_p, C.c_int32], C.c_int),
    "mpd_batch_tile": ([C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_double)], C.c_int),
    "mpd_init_ac_drop": ([C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int32, C.c_double,
                          C.c_double, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "mpd_init_ac_interface": ([C.c_void_p, C.c_void_p, C.c_int32, C.c_double, C.c_void_p, C.c_void_p,
                               C.c_void_p], C.c_int),
    "mpd_init_pp_drop": ([C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                          C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
                         C.c_int),
    "mpd_init_pp_flat": ([C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                          C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "mpd_init_pp_noise": ([C.c_void_p, C.c_double, C.c_double, C.c_uint64, C.c_double, C.c_void_p,
                           C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "mpd_device_init_drops": ([C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_double,
                               C.c_double], C.c_int),
    "mpd_device_init_interface": ([C.c_void_p, C.c_void_p, C.c_int32, C.c_double], C.c_int),
    "mpd_device_init_interface_noalloc": ([C.c_void_p, C.c_void_p, C.c_int32, C.c_double], C.c_int),
    "mpd_device_init_pp_noise": ([C.c_void_p, C.c_double, C.c_double, C.c_uint64, C.c_double], C.c_int),
    "mpd_device_init_pp_drop": ([C.c_void_p] + [C.c_double] * 7, C.c_int),
    "mpd_device_init_pp_flat": ([C.c_void_p] + [C.c_double] * 6, C.c_int),
    "mpd_default_output_info": ([C.c_void_p, C.c_void_p], C.c_int),
    "mpd_output_name": ([C.c_int32, C.c_void_p, C.c_int64, C.c_double, C.c_char_p, C.c_int32], C.c_int),
    "mpd_write_flowfield": ([C.c_char_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "mpd_write_profile": ([C.c_char_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "mpd_append_convergence": ([C.c_char_p, C.c_double, C.c_double], C.c_int),
    "mpd_default_case_info": ([C.c_void_p, C.c_void_p], C.c_int),
    "mpd_write_case": ([C.c_char_p, C.c_void_p], C.c_int),
}


class MpdError(RuntimeError):
    pass


def lib():
    """Load libmpdugks_hip.so (built by ``mpdugks_amd.build``); raise if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MpdError("libmpdugks_hip.so not built: run `python -m mpdugks_amd.build` "
                           "(or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SYMBOLS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        msg = lib().mpd_last_error().decode(errors="replace")
        raise MpdError("%s failed (%s): %s" % (what, abi.STATUS.get(rc, rc), msg))


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


class State:
    """Host copy of the step state in the ABI layout."""

    def __init__(self, cfg, h=None, f=None, m=None, ffx=None, ffy=None):
        N = cfg.nx * cfg.ny
        self.cfg = cfg
        self.h = h if h is not None else (np.zeros((9, N)) if abi.has_h(cfg.model) else None)
        self.f = f if f is not None else np.zeros((abi.num_q(cfg.model), N))
        self.m = m if m is not None else np.zeros((NM, N))
        self.ffx = ffx if ffx is not None else np.zeros(2 * N)
        self.ffy = ffy if ffy is not None else np.zeros(2 * N)

    def field(self, name):
        return self.m[M[name]].reshape(self.cfg.ny, self.cfg.nx)

    def copy(self):
        c = lambda a: None if a is None else a.copy()
        return State(self.cfg, c(self.h), c(self.f), c(self.m), c(self.ffx), c(self.ffy))


def ac_drop(cfg, cx=None, cy=None, radius=None, sign=+1, p0=1.0, mu0=0.1):
    """Initial state of AC_Drop (Ark/Cpp/Preprocess.cpp:705-760); defaults are the
    reference's centre (64, 64) and radius 0.36*ChLength scaled to the mesh."""
    L = cfg.nx * cfg.dx
    cx = 0.5 * L if cx is None else cx
    cy = 0.5 * cfg.ny * cfg.dx if cy is None else cy
    radius = 0.36 * L if radius is None else radius
    s = State(cfg)
    _check(lib().mpd_init_ac_drop(C.byref(cfg), cx, cy, radius, int(sign), p0, mu0, _p(s.h), _p(s.f),
                                  _p(s.m)), "mpd_init_ac_drop")
    return s


def ac_interface(cfg, y0, sign=+1, p0=1.0):
    """Interface along x at per-column heights y0 (nx values)."""
    y0 = np.ascontiguousarray(y0, dtype=np.float64)
    assert y0.shape == (cfg.nx,)
    s = State(cfg)
    _check(lib().mpd_init_ac_interface(C.byref(cfg), _p(y0), int(sign), p0, _p(s.h), _p(s.f), _p(s.m)),
           "mpd_init_ac_interface")
    return s


def scmp_drop(cfg, cx=None, cy=None, radius=None, rho_l=2.4806E-1, rho_v=4.5435E-2, width=5.0, mu0=0.1):
    """Initial state of SCMP_Drop (Ark/Cpp/Preprocess.cpp:337-401)."""
    L = cfg.nx * cfg.dx
    cx = 0.5 * L if cx is None else cx
    cy = 0.5 * cfg.ny * cfg.dx if cy is None else cy
    radius = 0.25 * L if radius is None else radius
    s = State(cfg)
    _check(lib().mpd_init_pp_drop(C.byref(cfg), cx, cy, radius, rho_l, rho_v, width, mu0, _p(s.f),
                                  _p(s.m), _p(s.ffx), _p(s.ffy)), "mpd_init_pp_drop")
    return s


def scmp_flat(cfg, y_lo=None, y_hi=None, slope=0.4, rho_l=2.4806E-1, rho_v=4.5435E-2, mu0=0.1):
    """Initial state of SCMP_FlatInterface (Ark/Cpp/Preprocess.cpp:291-336): a
    liquid band between y_lo and y_hi (default a quarter and three quarters of
    the mesh width, the reference's ChLength)."""
    L = cfg.nx * cfg.dx
    y_lo = 0.25 * L if y_lo is None else y_lo
    y_hi = 0.75 * L if y_hi is None else y_hi
    s = State(cfg)
    _check(lib().mpd_init_pp_flat(C.byref(cfg), y_lo, y_hi, slope, rho_l, rho_v, mu0, _p(s.f), _p(s.m),
                                  _p(s.ffx), _p(s.ffy)), "mpd_init_pp_flat")
    return s


def pp_noise(cfg, rho0, amp=0.01, seed=0, mu0=0.1):
    s = State(cfg)
    _check(lib().mpd_init_pp_noise(C.byref(cfg), rho0, amp, int(seed), mu0, _p(s.f), _p(s.m),
                                   _p(s.ffx), _p(s.ffy)), "mpd_init_pp_noise")
    return s


class Solver:
    """One case resident on the current GPU: the drop-in for DUGKS2DSolver()'s loop."""

    def __init__(self, cfg, path=None):
        self.cfg = cfg
        self._h = C.c_void_p()
        _check(lib().mpd_create(C.byref(cfg), C.byref(self._h)), "mpd_create")
        if path is not None:
            self.set_path(path)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().mpd_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_path(self, path):
        _check(lib().mpd_set_path(self._h, int(path)), "mpd_set_path")

    def set_face_order(self, order):
        """Pseudopotential face force order (abi.MPD_FACE_ORDER_*): Jacobi, or the
        reference's one-thread order (FP64 D2Q9, CD or UW, periodic)."""
        _check(lib().mpd_set_face_order(self._h, int(order)), "mpd_set_face_order")

    def upload_face_psi(self, psi):
        p = _f64(psi)
        if p is None or p.size != 2 * self.cfg.nx * self.cfg.ny:
            raise MpdError("upload_face_psi: need 2*nx*ny values")
        _check(lib().mpd_upload_face_psi(self._h, _p(p)), "mpd_upload_face_psi")

    def download_face_psi(self):
        p = np.zeros(2 * self.cfg.nx * self.cfg.ny)
        _check(lib().mpd_download_face_psi(self._h, _p(p)), "mpd_download_face_psi")
        return p

    def set_tile(self, tile):
        """FP32-bulk Allen-Cahn tile shape (abi.MPD_TILE_*; MPD_TILE_AUTO chooses per mesh)."""
        _check(lib().mpd_set_tile(self._h, int(tile)), "mpd_set_tile")

    def tile(self):
        """The tile shape (abi.MPD_TILE_*) the next step uses."""
        r = C.c_int32()
        _check(lib().mpd_get_tile(self._h, C.byref(r)), "mpd_get_tile")
        return r.value

    def init_drops(self, drops, sign=-1, p0=1.0, mu0=0.1):
        """Start state built on the GPU (mpd_device_init_drops): drops = [(cx, cy, r), ...]."""
        d = np.ascontiguousarray(drops, dtype=np.float64).reshape(-1, 3)
        cx, cy, r = (np.ascontiguousarray(d[:, k]) for k in range(3))
        _check(lib().mpd_device_init_drops(self._h, len(d), _p(cx), _p(cy), _p(r), int(sign), p0, mu0),
               "mpd_device_init_drops")

    def init_interface(self, y0, sign=+1, p0=1.0):
        y = self._heights(y0)
        _check(lib().mpd_device_init_interface(self._h, _p(y), int(sign), p0), "mpd_device_init_interface")

    def init_interface_noalloc(self, y0, sign=+1, p0=1.0):
        """init_interface through the entry point that allocates no device memory."""
        y = self._heights(y0)
        _check(lib().mpd_device_init_interface_noalloc(self._h, _p(y), int(sign), p0),
               "mpd_device_init_interface_noalloc")

    def _heights(self, y0):
        y = np.ascontiguousarray(y0, dtype=np.float64)
        if y.size != self.cfg.nx:
            raise MpdError("interface heights: need nx = %d values, got %d" % (self.cfg.nx, y.size))
        return y

    def init_pp_noise(self, rho0, amp=0.01, seed=0, mu0=0.1):
        _check(lib().mpd_device_init_pp_noise(self._h, rho0, amp, int(seed), mu0), "mpd_device_init_pp_noise")

    def init_pp_drop(self, cx=None, cy=None, radius=None, rho_l=2.4806E-1, rho_v=4.5435E-2, width=5.0, mu0=0.1):
        """SCMP_Drop built on the device (defaults as scmp_drop)."""
        L = self.cfg.nx * self.cfg.dx
        cx = 0.5 * L if cx is None else cx
        cy = 0.5 * self.cfg.ny * self.cfg.dx if cy is None else cy
        radius = 0.25 * L if radius is None else radius
        _check(lib().mpd_device_init_pp_drop(self._h, cx, cy, radius, rho_l, rho_v, width, mu0),
               "mpd_device_init_pp_drop")

    def init_pp_flat(self, y_lo=None, y_hi=None, slope=0.4, rho_l=2.4806E-1, rho_v=4.5435E-2, mu0=0.1):
        """SCMP_FlatInterface built on the device (defaults as scmp_flat)."""
        L = self.cfg.nx * self.cfg.dx
        y_lo = 0.25 * L if y_lo is None else y_lo
        y_hi = 0.75 * L if y_hi is None else y_hi
        _check(lib().mpd_device_init_pp_flat(self._h, y_lo, y_hi, slope, rho_l, rho_v, mu0),
               "mpd_device_init_pp_flat")

    def upload(self, s):
        # the library reads this many values from each array it takes
        N = self.cfg.nx * self.cfg.ny
        pp = not abi.has_h(self.cfg.model)
        need = {"h": None if pp else 9 * N, "f": abi.num_q(self.cfg.model) * N, "m": NM * N,
                "ffx": 2 * N if pp else None, "ffy": 2 * N if pp else None}
        arrs = []
        for name in ("h", "f", "m", "ffx", "ffy"):
            a = _f64(getattr(s, name))
            if a is not None and need[name] is not None and a.size != need[name]:
                raise MpdError("upload: %s holds %d values; this handle's %dx%d mesh needs %d"
                               % (name, a.size, self.cfg.nx, self.cfg.ny, need[name]))
            arrs.append(a)
        _check(lib().mpd_upload_state(self._h, *[_p(a) for a in arrs]), "mpd_upload_state")

    def download(self):
        s = State(self.cfg)
        _check(lib().mpd_download_state(self._h, _p(s.h), _p(s.f), _p(s.m), _p(s.ffx), _p(s.ffy)),
               "mpd_download_state")
        return s

    def step(self, n=1):
        _check(lib().mpd_step(self._h, int(n)), "mpd_step")

    def synchronize(self):
        _check(lib().mpd_synchronize(self._h), "mpd_synchronize")

    def reductions(self, update_residual=False):
        out = MpdSums()
        rc = lib().mpd_reductions(self._h, int(bool(update_residual)), C.byref(out))
        if rc not in (0, -3):
            _check(rc, "mpd_reductions")
        d = {k: getattr(out, k) for k, _ in MpdSums._fields_}
        d["status"] = rc
        return d

    def device_bytes(self):
        b = C.c_int64()
        _check(lib().mpd_device_bytes(self._h, C.byref(b)), "mpd_device_bytes")
        return b.value

    def set_kernel_timing(self, enable):
        _check(lib().mpd_set_kernel_timing(self._h, int(bool(enable))), "mpd_set_kernel_timing")

    def kernel_timing(self):
        ms, n = C.c_double(), C.c_int64()
        _check(lib().mpd_kernel_timing(self._h, C.byref(ms), C.byref(n)), "mpd_kernel_timing")
        return ms.value, n.value


class Batch:
    """Many FP32-bulk cases advanced by one kernel launch per step (mpd_batch_*)."""

    def __init__(self):
        self._b = C.c_void_p()
        _check(lib().mpd_batch_create(C.byref(self._b)), "mpd_batch_create")
        self.members = []

    def add(self, solver_):
        _check(lib().mpd_batch_add(self._b, solver_._h), "mpd_batch_add")
        self.members.append(solver_)

    def remove(self, solver_):
        _check(lib().mpd_batch_remove(self._b, solver_._h), "mpd_batch_remove")
        self.members.remove(solver_)

    def step(self, n):
        _check(lib().mpd_batch_step(self._b, int(n)), "mpd_batch_step")

    def size(self):
        nc, ncell = C.c_int32(), C.c_int64()
        _check(lib().mpd_batch_size(self._b, C.byref(nc), C.byref(ncell)), "mpd_batch_size")
        return nc.value, ncell.value

    def timing(self):
        ms, n = C.c_double(), C.c_int64()
        _check(lib().mpd_batch_timing(self._b, C.byref(ms), C.byref(n)), "mpd_batch_timing")
        return ms.value, n.value

    def set_tile(self, tile):
        """abi.MPD_TILE_AUTO: tuned per membership (default); another shape: forced."""
        _check(lib().mpd_batch_set_tile(self._b, int(tile)), "mpd_batch_set_tile")

    def tile(self):
        """(shape in use, [ms/step of 16x15, 16x16, 32x16] from the last tuning, 0 if not measured)."""
        r, ms = C.c_int32(), (C.c_double * 3)()
        _check(lib().mpd_batch_tile(self._b, C.byref(r), ms), "mpd_batch_tile")
        return r.value, list(ms)

    def close(self):
        if getattr(self, "_b", None) and self._b.value:
            lib().mpd_batch_destroy(self._b)
            self._b = C.c_void_p()
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_config(model=abi.MPD_MODEL_AC_HCZ):
    """The library's own defaults (mpd_default_config)."""
    cfg = MpdConfig()
    _check(lib().mpd_default_config(int(model), C.byref(cfg)), "mpd_default_config")
    return cfg


# ---------------- output files (Output.cpp formats; host only) ----------------

def output_info(cfg):
    """The reference's naming constants for cfg (mpd_default_output_info)."""
    o = abi.MpdOutputInfo()
    _check(lib().mpd_default_output_info(C.byref(cfg), C.byref(o)), "mpd_default_output_info")
    return o


def output_name(kind, info, step, t):
    buf = C.create_string_buffer(512)
    _check(lib().mpd_output_name(int(kind), C.byref(info), int(step), float(t), buf, 512), "mpd_output_name")
    return buf.value.decode()


def write_flowfield(path, cfg, state, info, order=None):
    m = _f64(state.m)
    o = None if order is None else np.ascontiguousarray(order, dtype=np.int64)
    _check(lib().mpd_write_flowfield(os.fsencode(path), cfg.nx, cfg.ny, _p(m), _p(o), C.byref(info)),
           "mpd_write_flowfield")


def write_profile(path, kind, cfg, state, info):
    m = _f64(state.m)
    _check(lib().mpd_write_profile(os.fsencode(path), int(kind), C.byref(cfg), _p(m), C.byref(info)),
           "mpd_write_profile")


def append_convergence(path, t, value):
    _check(lib().mpd_append_convergence(os.fsencode(path), float(t), float(value)), "mpd_append_convergence")


def case_info(cfg):
    c = abi.MpdCaseInfo()
    _check(lib().mpd_default_case_info(C.byref(cfg), C.byref(c)), "mpd_default_case_info")
    return c


def write_case(path, info):
    _check(lib().mpd_write_case(os.fsencode(path), C.byref(info)), "mpd_write_case")
