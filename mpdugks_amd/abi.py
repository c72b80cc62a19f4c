"""ctypes mirror of include/mpdugks.h (structs, enums) - no device code here."""
import ctypes as C

MPD_MODEL_AC_P, MPD_MODEL_AC_HCZ, MPD_MODEL_PP_D2Q9, MPD_MODEL_PP_D2V16 = 0, 1, 2, 3
MPD_FLUX_CD, MPD_FLUX_UW = 0, 1
MPD_BC_PERIODIC, MPD_BC_WALL_BB, MPD_BC_WALL_NEE = 0, 1, 2
MPD_FP64, MPD_FP32_BULK = 0, 1

MACRO_FIELDS = ["phi", "rho", "u", "v", "p", "phix", "phiy", "rhox", "rhoy", "fx", "fy",
                "lap", "ppu", "ppv", "psi", "scg", "mu", "tau"]
M = {name: k for k, name in enumerate(MACRO_FIELDS)}
NM = len(MACRO_FIELDS)

MODEL_NAMES = {"ac_p": MPD_MODEL_AC_P, "ac_hcz": MPD_MODEL_AC_HCZ,
               "pp_d2q9": MPD_MODEL_PP_D2Q9, "pp_d2v16": MPD_MODEL_PP_D2V16}

STATUS = {0: "MPD_OK", -1: "MPD_ERR_INVALID", -2: "MPD_ERR_DEVICE", -3: "MPD_ERR_NONFINITE",
          -4: "MPD_ERR_NOMEM", -5: "MPD_ERR_STATE"}


class MpdConfig(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("model", C.c_int32),
                ("flux", C.c_int32), ("bc_y", C.c_int32), ("precision", C.c_int32),
                ("dx", C.c_double), ("dt", C.c_double), ("rt", C.c_double),
                ("phi_l", C.c_double), ("phi_v", C.c_double),
                ("rho_l", C.c_double), ("rho_v", C.c_double),
                ("nu_l", C.c_double), ("nu_v", C.c_double),
                ("mobility", C.c_double), ("width", C.c_double), ("sigma", C.c_double),
                ("tau0", C.c_double), ("gx", C.c_double), ("gy", C.c_double),
                ("rho_ref", C.c_double), ("pp_tr", C.c_double), ("pp_k", C.c_double),
                ("pp_tau", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class MpdSums(C.Structure):
    _fields_ = [("sum_rho", C.c_double), ("sum_rho_u", C.c_double), ("sum_rho_v", C.c_double),
                ("sum_phi", C.c_double), ("kinetic_energy", C.c_double),
                ("max_speed", C.c_double), ("residual", C.c_double),
                ("nonfinite", C.c_int64)]


def num_q(model):
    return 16 if model == MPD_MODEL_PP_D2V16 else 9


def has_h(model):
    return model in (MPD_MODEL_AC_P, MPD_MODEL_AC_HCZ)

MPD_PATH_AUTO, MPD_PATH_PHASES, MPD_PATH_FUSED = 0, 1, 2
MPD_FACE_ORDER_JACOBI, MPD_FACE_ORDER_REFERENCE = 0, 1
MPD_TILE_AUTO, MPD_TILE_16X15, MPD_TILE_16X16, MPD_TILE_32X16 = 0, 1, 2, 3
TILE_NAMES = {MPD_TILE_16X15: "16x15", MPD_TILE_16X16: "16x16", MPD_TILE_32X16: "32x16"}

MPD_OUT_FLOWFIELD, MPD_OUT_SUMRHO, MPD_OUT_RESIDUAL, MPD_OUT_MIDX, MPD_OUT_MIDY, MPD_OUT_DIAG = range(6)


class MpdOutputInfo(C.Structure):
    _fields_ = [("ma", C.c_double), ("cfl", C.c_double), ("mu0", C.c_double), ("re", C.c_double),
                ("m_phi", C.c_double), ("rho_l", C.c_double), ("rho_v", C.c_double),
                ("temperature", C.c_double), ("mesh_nl", C.c_int32), ("mesh_type", C.c_char * 32),
                ("mesh_file", C.c_char * 128)]


class MpdCaseInfo(C.Structure):
    _fields_ = ([(k, C.c_char * n) for k, n in (("case_name", 64), ("mesh_file", 128), ("mesh_type", 32),
                                                ("flux_scheme", 16), ("bc", 32), ("velocity_model", 16),
                                                ("force_model", 32), ("eos_model", 32))]
                + [(k, C.c_int32) for k in ("zone_left", "zone_right", "zone_top", "zone_bottom")]
                + [(k, C.c_double) for k in ("omega0", "prandtl", "nk", "cv", "gamma")]
                + [(k, C.c_int32) for k in ("qu", "qv", "mesh_nl")]
                + [(k, C.c_double) for k in ("ch_length", "min_l", "kn", "ma", "re", "u0", "r0", "t0", "mu0",
                                             "nu0", "tau0", "lambda0", "cfl", "dt", "residual")]
                + [("end_step", C.c_int32)]
                + [(k, C.c_double) for k in ("maspan", "eta")]
                + [("phase_field", C.c_int32)]
                + [(k, C.c_double) for k in ("cn", "pe", "eo", "mo", "re_mp", "mobility", "tau_mass", "width",
                                             "sigma", "beta", "kappa", "rho_l", "rho_v", "mu_l", "mu_v", "nu_l",
                                             "nu_v", "radius")])
