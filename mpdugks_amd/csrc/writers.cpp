// Output files of the reference driver, byte for byte (Ark/Cpp/Output.cpp):
// the flow-field dump, the Convergence records and profiles, and Case.ark.  A
// host driver that steps on the GPU writes the same files from downloaded
// macro planes (structured order n = j*nx + i) as the reference writes from
// its CellArray.  Host code; iostream formatting is the reference's.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>

#include "../../include/mpdugks.h"
#include "last_error.h"

namespace {

const char *const FS = "    ";  // field separator of the reference's writers (ZeroMacros.h:21)

// a fixed-size text field of a caller's struct, read up to its first NUL or its
// end (a field that fills its array has no terminator)
template <size_t N>
std::string str(const char (&a)[N])
{
    return std::string(a, strnlen(a, N));
}

// MPD_ERR_INVALID with the reason for mpd_last_error
int bad(const char *fn, const char *what) { return mpd::set_error(MPD_ERR_INVALID, "%s: %s", fn, what); }
int cannot_open(const char *fn, const char *path)
{
    return mpd::set_error(MPD_ERR_INVALID, "%s: cannot open %s for writing", fn, path);
}
int written(const char *fn, const char *path, const std::ofstream &f)
{
    return f ? MPD_OK : mpd::set_error(MPD_ERR_INVALID, "%s: writing %s failed", fn, path);
}

int put_name(const std::string &s, char *buf, int32_t len)
{
    if (!buf || len <= 0) return bad("mpd_output_name", "null or empty output buffer");
    if ((int64_t)s.size() + 1 > len)
        return mpd::set_error(MPD_ERR_INVALID, "mpd_output_name: the name takes %zu bytes, the buffer %d",
                              s.size() + 1, len);
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return MPD_OK;
}

std::string tag(const mpd_output_info &o)
{
    // "Ma<Ma>_<mesh type><NL>_CFL<CFL>" (Output.cpp:319-320,487-488)
    std::ostringstream s;
    s << "Ma" << o.ma << "_" << str(o.mesh_type) << o.mesh_nl << "_CFL" << o.cfl;
    return s.str();
}

// thermodynamic pressure of the phase-field model (thermoPressure, Output.cpp:146-162)
double thermo_pressure(const mpd_config &c, const double *m, size_t N, size_t n)
{
    const double aphi = (c.rho_l - c.rho_v) / (c.phi_l - c.phi_v);
    const double beta = 12.0 * c.sigma / c.width, kappa = 3.0 * c.sigma * c.width / 2;
    const double phi = m[MPD_M_PHI * N + n];
    const double rx = m[MPD_M_RHOX * N + n], ry = m[MPD_M_RHOY * N + n];
    const double sq = (rx * rx + ry * ry) / (aphi * aphi);
    const double p0 = phi * 2 * beta * (phi - c.phi_l) * (phi - c.phi_v) * (2 * phi - c.phi_l - c.phi_v) -
                      beta * (phi - c.phi_l) * (phi - c.phi_l) * (phi - c.phi_v) * (phi - c.phi_v);
    return (p0 - kappa * phi * m[MPD_M_LAP * N + n] + kappa * sq / 2 + m[MPD_M_P * N + n]);
}

void copy_str(char *dst, size_t cap, const std::string &v)
{
    std::strncpy(dst, v.c_str(), cap - 1);
    dst[cap - 1] = 0;
}

double maxu(const mpd_config &c)
{
    // MaxU of the velocity set: sqrt(3 RT) for D2Q9 (D2Q9.h:16), 2 MaSpan sqrt(2) for D2V16 (D2V16.h:18-20)
    return c.model == MPD_MODEL_PP_D2V16 ? 2 * 1.2 * std::sqrt(2.0) : std::sqrt(3.0 * c.rt);
}

}  // namespace

extern "C" {

int mpd_default_output_info(const mpd_config *cfg, mpd_output_info *o)
{
    if (!cfg || !o) return bad("mpd_default_output_info", "null argument");
    std::memset(o, 0, sizeof(*o));
    const mpd_config &c = *cfg;
    o->ma = 0.0;                             // U0 / sqrt(R0 T0) with U0 = 0 (ZeroReference.h)
    o->cfl = c.dt * maxu(c) / c.dx;          // dt = CFL MinL / MaxU (ZeroNamespace.h:10-12)
    o->mu0 = c.tau0 * c.rt;                  // Mu0 = Rho0 Nu0, Tau0 = Nu0 / RT, Rho0 = 1
    o->re = 0.0;
    o->m_phi = c.mobility;
    o->rho_l = c.rho_l;
    o->rho_v = c.rho_v;
    o->temperature = c.rt;                   // T0 with R0 = 1
    o->mesh_nl = c.nx;
    copy_str(o->mesh_type, sizeof(o->mesh_type), "Car");
    copy_str(o->mesh_file, sizeof(o->mesh_file),
             std::to_string(c.nx) + "_" + std::to_string(c.ny) + "_Car_Periodic_Square_LBM");
    return MPD_OK;
}

int mpd_default_case_info(const mpd_config *cfg, mpd_case_info *ci)
{
    if (!cfg || !ci) return bad("mpd_default_case_info", "null argument");
    std::memset(ci, 0, sizeof(*ci));
    const mpd_config &c = *cfg;
    const bool ac = c.model == MPD_MODEL_AC_P || c.model == MPD_MODEL_AC_HCZ;
    const bool v16 = c.model == MPD_MODEL_PP_D2V16;
    copy_str(ci->case_name, sizeof(ci->case_name), ac ? "AC_Drop" : "SCMP_Drop");
    copy_str(ci->mesh_file, sizeof(ci->mesh_file),
             std::to_string(c.nx) + "_" + std::to_string(c.ny) + "_Car_Periodic_Square_LBM");
    copy_str(ci->mesh_type, sizeof(ci->mesh_type), "Car");
    copy_str(ci->flux_scheme, sizeof(ci->flux_scheme), c.flux == MPD_FLUX_UW ? "UW" : "CD");
    copy_str(ci->bc, sizeof(ci->bc), "Periodic");  // _BC_ARK as shipped (ZeroFlip.h:39-41)
    // DmQnName of each model's translation unit
    const char *vm = c.model == MPD_MODEL_AC_P ? "D2Q9AC" : c.model == MPD_MODEL_AC_HCZ ? "D2Q9Fak"
                     : c.model == MPD_MODEL_PP_D2Q9 ? "D2Q9PP" : "D2Q16";
    copy_str(ci->velocity_model, sizeof(ci->velocity_model), vm);
    copy_str(ci->force_model, sizeof(ci->force_model), "He-Shan-Doolean");
    copy_str(ci->eos_model, sizeof(ci->eos_model), "null");
    ci->zone_left = 4;
    ci->zone_right = 5;
    ci->zone_top = 7;
    ci->zone_bottom = 6;
    const double nk = 0.0;
    ci->omega0 = 0.5;
    ci->prandtl = 2.0 / 3.0;
    ci->nk = nk;
    ci->cv = (nk + 3.0) / 2.0;
    ci->gamma = (nk + 5.0) / (nk + 3.0);
    ci->qu = 1;
    ci->qv = v16 ? 16 : 9;
    ci->mesh_nl = c.nx;
    ci->ch_length = c.nx * c.dx;
    ci->min_l = c.dx;
    const double r0 = 1.0, t0 = c.rt / r0, rho0 = 1.0, lambda0 = 1 / (2.0 * r0 * t0);
    const double nu0 = c.tau0 * c.rt, mu0 = nu0 * rho0;
    ci->kn = 16.0 * mu0 / (5.0 * rho0 * r0 * t0) * std::sqrt(1.0 / (4.0 * lambda0 * M_PI));
    ci->ma = 0.0;
    ci->re = 0.0;
    ci->u0 = 0.0;
    ci->r0 = r0;
    ci->t0 = t0;
    ci->mu0 = mu0;
    ci->nu0 = nu0;
    ci->tau0 = c.tau0;
    ci->lambda0 = lambda0;
    ci->cfl = c.dt * maxu(c) / c.dx;
    ci->dt = c.dt;
    ci->residual = 1e-12;
    // End_Step = PhaseFieldAC::iT + 200 with iT = int(2 ChLength / U0 / dt): U0 = 0
    // makes it the x86 conversion of infinity, INT32_MIN (ZeroNamespace.h:128-133,169)
    ci->end_step = INT32_MIN + 200;
    ci->maspan = v16 ? 1.2 : std::sqrt(3.0 * c.rt);
    ci->eta = v16 ? 3.0 * 1.2 : 0.0;
    ci->phase_field = ac ? 1 : 0;
    ci->cn = c.width / ci->ch_length;
    ci->mobility = c.mobility;
    ci->tau_mass = c.mobility / c.rt;
    ci->width = c.width;
    ci->sigma = c.sigma;
    ci->beta = 12.0 * c.sigma / c.width;
    ci->kappa = 3.0 * c.sigma * c.width / 2;
    ci->rho_l = c.rho_l;
    ci->rho_v = c.rho_v;
    ci->mu_l = c.nu_l * c.rho_l;
    ci->mu_v = c.nu_v * c.rho_v;
    ci->nu_l = c.nu_l;
    ci->nu_v = c.nu_v;
    ci->radius = 0.36;  // radius / ChLength of AC_Drop (ZeroNamespace.h:113)
    return MPD_OK;
}

int mpd_output_name(int32_t kind, const mpd_output_info *o, int64_t step, double t, char *buf, int32_t len)
{
    if (!o) return bad("mpd_output_name", "null output info");
    std::ostringstream s;
    switch (kind) {
    case MPD_OUT_FLOWFIELD:  // Output_Flowfield, Output.cpp:486-488
        s << "global/" << "step" << step << tag(*o) << "_T" << t << ".dat";
        break;
    case MPD_OUT_SUMRHO:     // Output_SumRho, :258-261
        s << "Convergence/SumRho_mu" << o->mu0 << "_Re" << o->re << "_" << str(o->mesh_file) << ".dat";
        break;
    case MPD_OUT_RESIDUAL:   // Output_Residual, :272-275
        s << "Convergence/Residual_mu" << o->mu0 << "_Re" << o->re << "_" << str(o->mesh_file) << ".dat";
        break;
    case MPD_OUT_MIDX:       // Output_MidX, :318-320
    case MPD_OUT_MIDY:       // Output_MidY, :411-413
    case MPD_OUT_DIAG:       // Output_Diag, :464-466
        s << "Convergence/" << (kind == MPD_OUT_MIDX ? "MidX_" : kind == MPD_OUT_MIDY ? "MidY_" : "Diag_")
          << tag(*o) << "ratio_" << o->rho_l / o->rho_v << "step" << step << ".dat";
        break;
    default:
        return mpd::set_error(MPD_ERR_INVALID, "mpd_output_name: unknown kind %d", kind);
    }
    return put_name(s.str(), buf, len);
}

// Output_Flowfield (Output.cpp:484-555): zones Rho (101), U V (2), p (1), T (3),
// Phi (700), f.tau (701), Fx (702), Fy (703) over the cells in CellArray order
// (`order[k]` = structured index of the k-th cell; identity when null).
int mpd_write_flowfield(const char *path, int32_t nx, int32_t ny, const double *macro, const int64_t *order,
                        const mpd_output_info *o)
{
    if (!path || !macro || !o || nx <= 0 || ny <= 0)
        return bad("mpd_write_flowfield", "null argument or empty mesh");
    std::ofstream f(path);
    if (!f) return cannot_open("mpd_write_flowfield", path);
    f << std::setiosflags(std::ios::scientific) << std::setprecision(12);
    const size_t N = (size_t)nx * ny;
    auto cell = [&](size_t k) { return order ? (size_t)order[k] : k; };
    auto head = [&](int zone, int nd) { f << "(300 (" << zone << " 2 " << nd << " 0 0 1 " << N << ")(" << std::endl; };
    auto zone = [&](int id, int fld) {
        head(id, 1);
        for (size_t k = 0; k < N; ++k) f << macro[(size_t)fld * N + cell(k)] << std::endl;
        f << "))" << std::endl;
    };
    zone(101, MPD_M_RHO);
    head(2, 2);
    for (size_t k = 0; k < N; ++k)
        f << macro[MPD_M_U * N + cell(k)] << "  " << macro[MPD_M_V * N + cell(k)] << std::endl;
    f << "))" << std::endl;
    zone(1, MPD_M_P);
    head(3, 1);
    for (size_t k = 0; k < N; ++k) f << o->temperature << std::endl;
    f << "))" << std::endl;
    zone(700, MPD_M_PHI);
    zone(701, MPD_M_TAU);
    zone(702, MPD_M_FX);
    zone(703, MPD_M_FY);
    f.close();
    return written("mpd_write_flowfield", path, f);
}

// Output_MidX / Output_MidY / Output_Diag (Output.cpp:309-460): Rho, DP (p),
// TP (thermodynamic pressure), U, V along the column i = nx/2 - 1, the row
// j = ny/2 - 1, or the anti-diagonal (i, nx - 1 - i) (needs ny >= nx).
int mpd_write_profile(const char *path, int32_t kind, const mpd_config *cfg, const double *macro,
                      const mpd_output_info *o)
{
    if (!path || !cfg || !macro || !o) return bad("mpd_write_profile", "null argument");
    const int nx = cfg->nx, ny = cfg->ny;
    if (kind != MPD_OUT_MIDX && kind != MPD_OUT_MIDY && kind != MPD_OUT_DIAG)
        return mpd::set_error(MPD_ERR_INVALID, "mpd_write_profile: kind %d is not a profile", kind);
    if (kind == MPD_OUT_DIAG && ny < nx) return bad("mpd_write_profile", "the diagonal profile needs ny >= nx");
    std::ofstream f(path);
    if (!f) return cannot_open("mpd_write_profile", path);
    f << std::setiosflags(std::ios::scientific) << std::setprecision(kind == MPD_OUT_DIAG ? 8 : 20);
    const size_t N = (size_t)nx * ny;
    const int count = kind == MPD_OUT_MIDY || kind == MPD_OUT_DIAG ? nx : ny;
    const char *axis = kind == MPD_OUT_MIDY ? "X" : "Y";
    const int rows = kind == MPD_OUT_MIDY ? nx : ny;  // the head's "I = " (writeMidXHead/writeMidYHead)
    const char *names[5] = {"Rho", "DP", "TP", "U", "V"};
    for (int z = 0; z < 5; ++z) {
        f << "variables = " << axis << "," << names[z] << '\n'
          << "zone T = " << names[z] << "_M" << o->m_phi << " " << "I = " << rows << std::endl;
        for (int k = 0; k < count; ++k) {
            int i, j;
            if (kind == MPD_OUT_MIDX) { i = nx / 2 - 1; j = k; }
            else if (kind == MPD_OUT_MIDY) { i = k; j = ny / 2 - 1; }
            else { i = k; j = nx - 1 - k; }
            const size_t n = (size_t)j * nx + i;
            const double pos = kind == MPD_OUT_MIDX ? (j + 0.5) * cfg->dx : (i + 0.5) * cfg->dx;
            double v;
            switch (z) {
            case 0: v = macro[MPD_M_RHO * N + n]; break;
            case 1: v = macro[MPD_M_P * N + n]; break;
            case 2: v = thermo_pressure(*cfg, macro, N, n); break;
            case 3: v = macro[MPD_M_U * N + n]; break;
            default: v = macro[MPD_M_V * N + n]; break;
            }
            f << pos << FS << v << std::endl;
        }
    }
    f.close();
    return written("mpd_write_profile", path, f);
}

// Output_SumRho / Output_Residual (Output.cpp:258-287): append "t    value".
int mpd_append_convergence(const char *path, double t, double value)
{
    if (!path) return bad("mpd_append_convergence", "null path");
    std::ofstream f(path, std::ofstream::app);
    if (!f) return cannot_open("mpd_append_convergence", path);
    f << std::setiosflags(std::ios::scientific) << std::setprecision(8);
    f << t << FS << value << std::endl;
    f.close();
    return written("mpd_append_convergence", path, f);
}

// OutputCase (Output.cpp:65-144): Case.ark.
int mpd_write_case(const char *path, const mpd_case_info *c)
{
    if (!path || !c) return bad("mpd_write_case", "null argument");
    std::ofstream f(path);
    if (!f) return cannot_open("mpd_write_case", path);
    const char *e = "    =    ";
    f << str(c->case_name) << e << "Case Name" << std::endl
      << str(c->mesh_file) << e << "Mesh File" << std::endl
      << str(c->mesh_type) << e << "Mesh Type" << std::endl
      << str(c->flux_scheme) << e << "Flux Scheme" << std::endl
      << str(c->bc) << e << "Boundary Condition" << std::endl
      << str(c->velocity_model) << e << "velocity model" << std::endl
      << str(c->force_model) << e << "Force model" << std::endl
      << str(c->eos_model) << e << "EoS model" << std::endl
      << c->zone_left << FS << "=" << FS << "left" << std::endl
      << c->zone_right << FS << "=" << FS << "right" << std::endl
      << c->zone_top << FS << "=" << FS << "top" << std::endl
      << c->zone_bottom << FS << "=" << FS << "bottom" << std::endl
      << "#--------------thermodynamics------------" << std::endl
      << c->omega0 << e << "Omega0//sutherland power" << std::endl
      << c->prandtl << e << "Prandtl" << std::endl
      << c->nk << e << "nK//internal degrees of freedom" << std::endl
      << c->cv << e << "specific heat of constant volume" << std::endl
      << c->gamma << e << "Gamma//specific heat ratio" << std::endl
      << "#----------------Reference----------------" << std::endl
      << c->qu << e << "Qu" << std::endl
      << c->qv << e << "Qv" << std::endl
      << c->mesh_nl << e << "NL" << std::endl
      << c->ch_length << e << "Characteristic Length" << std::endl
      << c->min_l << e << "MinL" << std::endl
      << c->kn << e << "Knudsen" << std::endl
      << c->ma << e << "Mach" << std::endl
      << c->re << e << "Re" << std::endl
      << c->u0 << e << "U0" << std::endl
      << c->r0 << e << "R0" << std::endl
      << c->t0 << e << "T0" << std::endl
      << c->mu0 << e << "dynamic viscosity" << std::endl
      << c->nu0 << e << "kinetic viscosity" << std::endl
      << c->tau0 << e << "tau0" << std::endl
      << c->lambda0 << e << "Lambda0" << std::endl
      << c->cfl << e << "CFL" << std::endl
      << c->dt << e << "dt" << std::endl
      << c->dt / c->tau0 << e << "dt/tau0" << std::endl
      << c->residual << e << "RESIDUAL" << std::endl
      << c->end_step << e << "End_Step" << std::endl
      << c->maspan << e << "MaSpan" << std::endl
      << c->eta << e << "Eta// T of D2V16" << std::endl;
    if (c->phase_field) {
        f << "-------------Phase Field--------------" << std::endl
          << c->cn << e << "Cahn number" << std::endl
          << c->pe << e << "Peclet number" << std::endl
          << c->eo << e << "Eotvos(Bond) number" << std::endl
          << c->mo << e << "Morton number" << std::endl
          << c->re_mp << e << "Reynolds number" << std::endl
          << std::endl
          << c->mobility << e << "mobility" << std::endl
          << c->tau_mass << e << "mobility relaxition time" << std::endl
          << c->dt / c->tau_mass << e << "dt/tauMass" << std::endl
          << c->width << e << "inteface width" << std::endl
          << c->sigma << e << "Sigma(surface tension coefficient)" << std::endl
          << c->beta << e << "Beta" << std::endl
          << c->kappa << e << "Kappa" << std::endl
          << c->rho_l << e << "liquid density" << std::endl
          << c->rho_v << e << "vapor density" << std::endl
          << c->mu_l << e << "liquid dynamic viscosity" << std::endl
          << c->mu_v << e << "vapor dynamic viscosity" << std::endl
          << c->nu_l << e << "liquid kinetic viscosity" << std::endl
          << c->nu_v << e << "vapor kinetic viscosity" << std::endl
          << c->rho_l / c->rho_v << e << "density ratio" << std::endl
          << c->mu_l / c->mu_v << e << "dynamic viscosity ratio" << std::endl
          << c->radius << e << "radius" << std::endl;
    }
    f << "-------------End--------------" << std::endl;
    f.close();
    return written("mpd_write_case", path, f);
}

}  // extern "C"
