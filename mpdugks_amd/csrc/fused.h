// Fused one-kernel-per-step path: derived state, tile kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mpdugks.h"
#include "device_state.h"

namespace mpd {

struct Consts;

// FP32-bulk pseudopotential D2Q9 planes (pp32.hip)
struct PP32Planes {
    int nx, ny;
    int64_t N;
    double *rho;                                  // [N] conserved density (FP64)
    float *f, *fbp;                               // [9][N]
    float *U, *V, *tau, *Fx, *Fy, *psi, *scg;     // [N]
    float *ffx, *ffy;                             // face force [2N]
    float *fpsi, *frho, *fU, *fV, *fscg, *fmass;  // face scratch [2N]
    float *ffl;                                   // face fluxes [9][2N]
};
// the carried planes of one state set: the one-launch step reads the current
// set (with its halos) and writes the other
struct PP32Set {
    double *rho;
    float *f, *U, *V, *Fx, *Fy, *psi, *scg, *ffx, *ffy;
};
struct PP32State {
    PP32Planes P;         // the current set, tau, and the scratch of the per-phase kernels
    PP32Set other;        // the second set
    bool phases = false;  // step with one launch per phase instead of the one-launch tile kernel
    int cus = 256;        // compute units of the device (tile shape choice)
    int tile = MPD_TILE_AUTO;  // tile shape of the one-launch step: 16x15, 32x16, or chosen per mesh
};
// tile shape (MPD_TILE_16X15 or MPD_TILE_32X16) of the FP32-bulk pseudopotential step
int pp32_tile_auto(int nx, int ny, int cus);
int pp32_alloc(PP32State &S, void **mem, int64_t *bytes, const Consts &C, int64_t N);
void pp32_from64(const PP32State &S, const DevState &D, hipStream_t st);
// force_stale: the stored cell force is the previous step's (every step but the
// first after a conversion); the step or the conversion recomputes it from Psi
void pp32_to64(const PP32State &S, const DevState &D, const Consts &C, bool force_stale, hipStream_t st);
void pp32_step(PP32State &S, const Consts &C, bool force_stale, hipStream_t st);

struct FusedState {
    bool ready = false;
    int64_t bytes = 0;
    void *mem[16] = {nullptr};
    int nmem = 0;
    // ping-pong state planes (element type depends on precision)
    void *st[2] = {nullptr, nullptr};
    void *gl[2] = {nullptr, nullptr};  // walls: lagged ghost phi rows [2][nx]
    double *phd[2] = {nullptr, nullptr};  // FP32-bulk: the conserved phi in FP64 [N]
    int cur = 0;
    int nplanes = 0;
    int64_t ncell = 0;
    bool fp32 = false;
    double p_ref = 0.0;  // FP32: f~ stored minus the rest equilibrium at p_ref
    bool pp = false;     // FP32-bulk pseudopotential: explicit FP32 planes, no derived state
    PP32State pp32;
    bool pp_force_stale = false;  // the stored cell force predates the last update
    int tile = MPD_TILE_16X15;   // Allen-Cahn: tile shape of the step kernel (mpd_tile)
};

// FP32 planes of the derived state (see fused_kernels.hip); phd is the
// conserved phase field in FP64, phi its FP32 copy for the stencils
struct FusedPlanes32 {
    float *h, *f, *phi, *phl, *ppu, *ppv, *gl;
    double *phd;
};
// one FP32-bulk step (fused_fp32.hip) with tiles of shape `tile` (mpd_tile)
void fused32_step(const FusedPlanes32 &in, const FusedPlanes32 &out, const Consts &C, int64_t N, int tile,
                  hipStream_t st);
// tile shape (mpd_tile) of the FP32-bulk step for an nx x ny mesh on a GPU with
// `cus` compute units, whose two state sets take `state_bytes` of device memory
int fused32_tile_auto(int nx, int ny, int cus, int64_t state_bytes);
// tile shape (MPD_TILE_16X15 or MPD_TILE_16X16) of the FP64 Allen-Cahn step
int fused64_tile_auto(int nx, int ny, int cus);
// device bytes of the two FP32-bulk Allen-Cahn state sets of a configuration
int64_t fused32_state_bytes(const mpd_config &cfg);
// planes of one Allen-Cahn state set (fused_alloc)
int fused_nplanes(const mpd_config &cfg);

// batched FP32-bulk step over many cases (fused_fp32.hip)
struct BatchMember {
    const Consts *C;
    int64_t N;
    FusedPlanes32 pl[2];
    int cur;
};
struct Batch32Plan;
Batch32Plan *batch32_new();
void batch32_free(Batch32Plan *P);
int batch32_plan(Batch32Plan *P, const BatchMember *m, int n, int tile, hipStream_t st);
// one step of every case; parity 0/1 selects in/out relative to each case's cur
void batch32_step(const Batch32Plan *P, int parity, hipStream_t st);
// FP32 planes of a fused state (set 0/1)
FusedPlanes32 fused32_planes(FusedState &F, int which);

bool fused_supported(const mpd_config &cfg);
// the path starts from two explicit FP64 steps (the Allen-Cahn derived state)
bool fused_needs_priming(const mpd_config &cfg);
int fused_alloc(FusedState &F, const DevState &S, const Consts &C, const mpd_config &cfg);
void fused_free(FusedState &F);
int64_t fused_bytes(const FusedState &F);
// first != 0: take the explicit macros of S as the state of this step
void fused_step(FusedState &F, const DevState &S, const Consts &C, const mpd_config &cfg, int first,
                hipStream_t st);
// priming from the explicit state S: which = 0 before the first phase step
// (walls: lagged ghost phi), 1 after it (phi_lag, prevPhiU/V), 2 after the
// second (populations, phi)
void fused_convert(FusedState &F, const DevState &S, const Consts &C, int which, hipStream_t st);
// write the full macro set (and tilde planes) of the current fused state into S
int fused_materialize(FusedState &F, const DevState &S, const Consts &C, const mpd_config &cfg,
                      hipStream_t st);

}  // namespace mpd
