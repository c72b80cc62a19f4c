// Device buffers of one case (structure-of-arrays planes, cell n = j*nx + i).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mpd {

struct DevState {
    int64_t N;        // cells
    // state carried by the reference between steps
    double *h, *f;    // [Q][N] tilde populations (h: Allen-Cahn, f: momentum)
    double *m;        // [MPD_M_COUNT][N] cell macros
    double *ffx, *ffy;// pseudopotential face force, x-faces [N] then y-faces [N]
    // scratch of the phase path
    double *hbp, *fbp;            // [Q][N] bar-plus
    double *hbpx, *hbpy, *fbpx, *fbpy; // [Q][N] UW gradients
    double *hfl, *ffl;            // face fluxes: x-faces [Q][N] then y-faces [Q][nyf*nx]
    double *gm, *ghbp, *gfbp;     // wall ghost rows
    double *cbhl, *cbfl;          // UW: fluxes of the second copies of the periodic faces
                                  // [Q][nx + ny] (x-faces by row, then y-faces by column)
    // pseudopotential D2Q9 with UW faces in the reference face order: the second
    // copies' Psi (this step, previous step), previous force and face macros [nx + ny]
    double *cbpsi, *cbpsi_old, *cbffx, *cbffy, *cbscr;
    double *fpsi, *fscr;          // PP face Psi [2N], face scratch [4][2N]
    double *fpsi_old;             // PP face Psi of the previous step [2N] (0 before the first step)
    double *rho1k;                // density of the last residual evaluation
};

struct Consts;
void phase_step(const DevState &S, const Consts &C, hipStream_t st);
// second copies of the periodic faces from the first copies (the face at i = 0
// of each row, j = 0 of each column): forces (ffx, ffy) and/or previous Psi
void pp_copies_from_first(const DevState &S, const Consts &C, bool forces, bool psi, hipStream_t st);

// start states built on the device (devinit.hip)
constexpr int MAX_DROPS = 8;
struct DropSet {
    int n;
    double cx[MAX_DROPS], cy[MAX_DROPS], r[MAX_DROPS];
    double width;
};
void devinit_drops(const DevState &S, const Consts &C, const DropSet &D, int sign, double p0, double mu0,
                   hipStream_t st);
void devinit_interface(const DevState &S, const Consts &C, const double *d_y0, int sign, double p0, double width,
                       hipStream_t st);
void devinit_noise(const DevState &S, const Consts &C, double rho0, double amp, uint64_t seed, double mu0,
                   hipStream_t st);
void devinit_pp_drop(const DevState &S, const Consts &C, double cx, double cy, double radius, double rho_l,
                     double rho_v, double width, double mu0, hipStream_t st);
void devinit_pp_flat(const DevState &S, const Consts &C, double y_lo, double y_hi, double slope, double rho_l,
                     double rho_v, double mu0, hipStream_t st);

}  // namespace mpd
