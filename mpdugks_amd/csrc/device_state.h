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
    double *fpsi, *fscr;          // PP face Psi [2N], face scratch [4][2N]
    double *rho1k;                // density of the last residual evaluation
};

struct Consts;
void phase_step(const DevState &S, const Consts &C, hipStream_t st);

}  // namespace mpd
