// Host-side initial states in the library's upload layout.  These restate the
// reference initialisers so that a driver (or a test) can build the state the
// time loop starts from without the reference's mesh objects.
#include <cmath>
#include <cstring>
#include <cstdint>
#include <random>

#include "../../include/mpdugks.h"
#include "d2v16.h"
#include "last_error.h"

namespace {

const double W9[9] = {4.0 / 9.0, 1.0 / 9.0, 1.0 / 36.0, 1.0 / 9.0, 1.0 / 36.0,
                      1.0 / 9.0, 1.0 / 36.0, 1.0 / 9.0, 1.0 / 36.0};
const int EX[9] = {0, 1, 1, 0, -1, -1, -1, 0, 1};
const int EY[9] = {0, 0, 1, 1, 1, 0, -1, -1, -1};

constexpr mpd::d2v16::Table kV16 = mpd::d2v16::make_table();

int num_q(const mpd_config &c) { return c.model == MPD_MODEL_PP_D2V16 ? mpd::d2v16::Q : 9; }

// Update_phi_Eq with zero sources for one cell (HCZ Ark/Lib/D2Q9HCZ.cpp:54-92,
// AC-p Ark/Lib/D2Q9AC.cpp:69-105, PP Ark/Lib/D2Q9PP.cpp:51-62; D2V16
// Ark/Lib/D2V16.cpp:157-166 with Macro16 at the reference temperature).
void cell_eq(const mpd_config &c, double phi, double rho, double U, double V, double p, double *heq,
             double *feq)
{
    if (c.model == MPD_MODEL_PP_D2V16) {
        double M[16];
        mpd::d2v16::macro16<double>(M, rho, U, V, c.rt, rho * c.rt);
        for (int j = 0; j < 16; ++j) {
            double s = 0;
            for (int k = 0; k < 16; ++k) s += kV16.a[j][k] * M[k];
            feq[j] = s;
        }
        return;
    }
    const double rt = c.rt, lambda0 = 1 / (2.0 * rt), maspan = std::sqrt(3.0 * rt);
    double uu = U * U + V * V;
    for (int q = 0; q < 9; ++q) {
        double ex = maspan * EX[q], ey = maspan * EY[q], w = W9[q];
        double u1 = (ex * U + ey * V) / rt;
        double G = u1 + 0.5 * u1 * u1 - uu * lambda0;
        if (c.model == MPD_MODEL_PP_D2Q9) {
            feq[q] = w * rho * (1.0 + u1 + 0.5 * u1 * u1 - uu * lambda0);
        } else {
            heq[q] = w * (1.0 + u1) * phi;
            if (c.model == MPD_MODEL_AC_HCZ) feq[q] = w * (p + rho * G * rt);
            else feq[q] = w * (p / rt + rho * G) - (double)(q == 0) * p / rt;
        }
    }
}

// 0, or MPD_ERR_INVALID with the reason for mpd_last_error: fn's config, its
// model family (Allen-Cahn, or pseudopotential) and its output arrays
int check(const char *fn, const mpd_config *c, bool allen_cahn, bool arrays)
{
    if (!c) return mpd::set_error(MPD_ERR_INVALID, "%s: null config", fn);
    if (c->model < MPD_MODEL_AC_P || c->model > MPD_MODEL_PP_D2V16)
        return mpd::set_error(MPD_ERR_INVALID, "%s: unknown model %d", fn, c->model);
    if (!(c->nx > 0 && c->ny > 0 && c->rt > 0))
        return mpd::set_error(MPD_ERR_INVALID, "%s: need nx, ny and rt > 0", fn);
    if (allen_cahn != (c->model < MPD_MODEL_PP_D2Q9))
        return mpd::set_error(MPD_ERR_INVALID, "%s: starts the %s models", fn,
                              allen_cahn ? "Allen-Cahn" : "pseudopotential");
    if (!arrays) return mpd::set_error(MPD_ERR_INVALID, "%s: null input or output array", fn);
    return 0;
}

}  // namespace

extern "C" {

// AC_Drop (Ark/Cpp/Preprocess.cpp:705-760) with AnalyticalPhiAC_Drop (:402-419):
// phi = (PhiL+PhiV)/2 + sign*(PhiL-PhiV)/2*tanh(2(r-R)/W), sign = +1 reproduces the
// reference (light phase inside, a bubble), sign = -1 gives a droplet.  The macros
// start from MacroQuantity(Rho0=1, U0=0, V0=0, p0=1, ..., Mu0) (ZeroReference.h:78-92)
// and f.tau = Mu0/(rho RT).  tilde = equilibrium.
int mpd_init_ac_drop(const mpd_config *cfg, double cx, double cy, double radius, int32_t sign,
                     double p0, double mu0, double *tilde_h, double *tilde_f, double *macro)
{
    if (int r = check("mpd_init_ac_drop", cfg, true, tilde_h && tilde_f && macro)) return r;
    const mpd_config &c = *cfg;
    const size_t N = (size_t)c.nx * c.ny;
    const double aphi = (c.rho_l - c.rho_v) / (c.phi_l - c.phi_v), bphi = c.rho_v - aphi * c.phi_v;
    std::memset(macro, 0, sizeof(double) * MPD_M_COUNT * N);
    double heq[9], feq[9];
    for (int j = 0; j < c.ny; ++j)
        for (int i = 0; i < c.nx; ++i) {
            size_t n = (size_t)j * c.nx + i;
            double x = (i + 0.5) * c.dx, y = (j + 0.5) * c.dx;
            double t = std::tanh((std::sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy)) - radius) * 2 / c.width);
            if (sign < 0) t = -t;
            double phi = 0.5 * (c.phi_l + c.phi_v) + 0.5 * (c.phi_l - c.phi_v) * t;
            double rho = aphi * phi + bphi;
            macro[MPD_M_PHI * N + n] = phi;
            macro[MPD_M_RHO * N + n] = rho;
            macro[MPD_M_P * N + n] = p0;
            macro[MPD_M_MU * N + n] = mu0;
            macro[MPD_M_TAU * N + n] = mu0 / (rho * c.rt);
            cell_eq(c, phi, rho, 0.0, 0.0, p0, heq, feq);
            for (int q = 0; q < 9; ++q) {
                tilde_h[q * N + n] = heq[q];
                tilde_f[q * N + n] = feq[q];
            }
        }
    return 0;
}

// Flat interface along x with per-column height y0[i] (Rayleigh-Taylor start;
// the reference has no such initialiser): phi = (PhiL+PhiV)/2 +
// sign*(PhiL-PhiV)/2*tanh(2(y-y0)/W) (sign = +1: phase L above), u = 0, p = p0,
// Mu = rho*(NuV + (phi-PhiV)(NuL-NuV)) (the HCZ viscosity, D2Q9HCZ.cpp:24-40),
// f.tau = Mu/(rho RT), tilde = equilibrium.
int mpd_init_ac_interface(const mpd_config *cfg, const double *y0, int32_t sign, double p0, double *tilde_h,
                          double *tilde_f, double *macro)
{
    if (int r = check("mpd_init_ac_interface", cfg, true, y0 && tilde_h && tilde_f && macro)) return r;
    const mpd_config &c = *cfg;
    const size_t N = (size_t)c.nx * c.ny;
    const double aphi = (c.rho_l - c.rho_v) / (c.phi_l - c.phi_v), bphi = c.rho_v - aphi * c.phi_v;
    std::memset(macro, 0, sizeof(double) * MPD_M_COUNT * N);
    double heq[9], feq[9];
    for (int j = 0; j < c.ny; ++j)
        for (int i = 0; i < c.nx; ++i) {
            size_t n = (size_t)j * c.nx + i;
            double y = (j + 0.5) * c.dx;
            double t = std::tanh(2.0 * (y - y0[i]) / c.width);
            if (sign < 0) t = -t;
            double phi = 0.5 * (c.phi_l + c.phi_v) + 0.5 * (c.phi_l - c.phi_v) * t;
            double rho = aphi * phi + bphi;
            double mu = (c.nu_v + (phi - c.phi_v) * (c.nu_l - c.nu_v)) * rho;
            macro[MPD_M_PHI * N + n] = phi;
            macro[MPD_M_RHO * N + n] = rho;
            macro[MPD_M_P * N + n] = p0;
            macro[MPD_M_MU * N + n] = mu;
            macro[MPD_M_TAU * N + n] = mu / (rho * c.rt);
            cell_eq(c, phi, rho, 0.0, 0.0, p0, heq, feq);
            for (int q = 0; q < 9; ++q) {
                tilde_h[q * N + n] = heq[q];
                tilde_f[q * N + n] = feq[q];
            }
        }
    return 0;
}

// SCMP_Drop (Ark/Cpp/Preprocess.cpp:337-401): rho = (RhoL+RhoV)/2 - (RhoV-RhoL)/2 *
// tanh(2(r-R)/width), p = rho RT, F = 0, tau = pp_tau, tilde = equilibrium; face
// forces are the owner/neighbour averages of the (zero) cell forces.
int mpd_init_pp_drop(const mpd_config *cfg, double cx, double cy, double radius, double rho_l, double rho_v,
                     double width, double mu0, double *tilde_f, double *macro, double *face_fx, double *face_fy)
{
    if (int r = check("mpd_init_pp_drop", cfg, false, tilde_f && macro)) return r;
    const mpd_config &c = *cfg;
    const size_t N = (size_t)c.nx * c.ny;
    std::memset(macro, 0, sizeof(double) * MPD_M_COUNT * N);
    double feq[16], heq[16];
    const int nq = num_q(c);
    for (int j = 0; j < c.ny; ++j)
        for (int i = 0; i < c.nx; ++i) {
            size_t n = (size_t)j * c.nx + i;
            double x = (i + 0.5) * c.dx, y = (j + 0.5) * c.dx;
            double rho = 0.5 * (rho_l + rho_v) - 0.5 * (-rho_l + rho_v) *
                         std::tanh(2 * (std::sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy)) - radius) / width);
            macro[MPD_M_RHO * N + n] = rho;
            macro[MPD_M_P * N + n] = rho * c.rt;
            macro[MPD_M_MU * N + n] = mu0;
            macro[MPD_M_TAU * N + n] = c.pp_tau;
            cell_eq(c, 0.0, rho, 0.0, 0.0, rho * c.rt, heq, feq);
            for (int q = 0; q < nq; ++q) tilde_f[q * N + n] = feq[q];
        }
    if (face_fx) std::memset(face_fx, 0, sizeof(double) * 2 * N);
    if (face_fy) std::memset(face_fy, 0, sizeof(double) * 2 * N);
    return 0;
}

// SCMP_FlatInterface (Ark/Cpp/Preprocess.cpp:291-336): a liquid band between
// y_lo and y_hi, rho = rho_v + (rho_l - rho_v)/2 (tanh(slope (y - y_lo)) -
// tanh(slope (y - y_hi))), p = rho RT, u = 0, tilde = equilibrium.  The
// reference's case is slope 0.4, y_lo = ChLength/4, y_hi = 3 ChLength/4.
int mpd_init_pp_flat(const mpd_config *cfg, double y_lo, double y_hi, double slope, double rho_l, double rho_v,
                     double mu0, double *tilde_f, double *macro, double *face_fx, double *face_fy)
{
    if (int r = check("mpd_init_pp_flat", cfg, false, tilde_f && macro)) return r;
    const mpd_config &c = *cfg;
    const size_t N = (size_t)c.nx * c.ny;
    std::memset(macro, 0, sizeof(double) * MPD_M_COUNT * N);
    double feq[16], heq[16];
    const int nq = num_q(c);
    for (int j = 0; j < c.ny; ++j) {
        const double y = (j + 0.5) * c.dx;
        const double rho = rho_v + 0.5 * (rho_l - rho_v) * (std::tanh(slope * (y - y_lo)) - std::tanh(slope * (y - y_hi)));
        cell_eq(c, 0.0, rho, 0.0, 0.0, rho * c.rt, heq, feq);
        for (int i = 0; i < c.nx; ++i) {
            const size_t n = (size_t)j * c.nx + i;
            macro[MPD_M_RHO * N + n] = rho;
            macro[MPD_M_P * N + n] = rho * c.rt;
            macro[MPD_M_MU * N + n] = mu0;
            macro[MPD_M_TAU * N + n] = c.pp_tau;
            for (int q = 0; q < nq; ++q) tilde_f[q * N + n] = feq[q];
        }
    }
    if (face_fx) std::memset(face_fx, 0, sizeof(double) * 2 * N);
    if (face_fy) std::memset(face_fy, 0, sizeof(double) * 2 * N);
    return 0;
}

// Liquid-vapour separation start (pseudopotential): rho = rho0*(1 + U(-amp, amp))
// per cell from a seeded std::mt19937_64, u = 0, tilde = equilibrium.
int mpd_init_pp_noise(const mpd_config *cfg, double rho0, double amp, uint64_t seed, double mu0,
                      double *tilde_f, double *macro, double *face_fx, double *face_fy)
{
    if (int r = check("mpd_init_pp_noise", cfg, false, tilde_f && macro)) return r;
    const mpd_config &c = *cfg;
    const size_t N = (size_t)c.nx * c.ny;
    std::memset(macro, 0, sizeof(double) * MPD_M_COUNT * N);
    std::mt19937_64 gen(seed);
    std::uniform_real_distribution<double> U(-amp, amp);
    double feq[16], heq[16];
    const int nq = num_q(c);
    for (size_t n = 0; n < N; ++n) {
        double rho = rho0 * (1.0 + U(gen));
        macro[MPD_M_RHO * N + n] = rho;
        macro[MPD_M_P * N + n] = rho * c.rt;
        macro[MPD_M_MU * N + n] = mu0;
        macro[MPD_M_TAU * N + n] = c.pp_tau;
        cell_eq(c, 0.0, rho, 0.0, 0.0, rho * c.rt, heq, feq);
        for (int q = 0; q < nq; ++q) tilde_f[q * N + n] = feq[q];
    }
    if (face_fx) std::memset(face_fx, 0, sizeof(double) * 2 * N);
    if (face_fy) std::memset(face_fy, 0, sizeof(double) * 2 * N);
    return 0;
}

}  // extern "C"
