// Fused one-kernel-per-step path: derived state, tile kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mpdugks.h"
#include "device_state.h"

namespace mpd {

struct Consts;

// BGK combinations a x + b x_eq + c x_src (a = 1 - b) with b folded into the
// equilibrium: (1 - b) x + (b x_eq) + c x_src.  From 1 - b rounded to FP32 (a),
// the FP32 pair (1 - b, b) whose sum is exactly 1: the member >= 1/2 keeps a
// rounded value and the other is its exact complement (Sterbenz), so the unit
// coefficient that x + b (x_eq - x) has exactly carries no rounding either.
__host__ __device__ __forceinline__ void split1(float a, float &omb, float &b)
{
    b = 1.f - a;
    omb = 1.f - b;
}

// FP32-bulk pseudopotential planes (pp32.hip), velocity set of q populations
// (D2Q9 or D2V16).  A state set is one float block of (q + 4) planes -- f~[q],
// U, V, and the face Psi of the step that produced it, signed by scG (x-faces,
// then y-faces) -- and the conserved density in FP64.  The shared planes hold
// what a step does not change (tau), the forces of an uploaded state (read by
// the first step after an upload) or of a download, and the scratch of the
// one-launch-per-phase kernels (D2Q9).
struct PP32Planes {
    int nx, ny, q;
    int64_t N;
    double *rho;                                  // current set: conserved density (FP64) [N]
    float *f;                                     // current set: float block [(q + 4) N]
    float *tau, *Fx, *Fy;                         // [N]: relaxation time, cell force
    float *ffx, *ffy;                             // [2N]: face force
    float *fbp, *psi;                             // per-phase scratch: bar-plus [9][N], cell scG Psi [N]
    float *fpsi, *frho, *fU, *fV, *fmass;         // per-phase face scratch: scG Psi, rho, U, V, mass [2N]
    float *ffl;                                   // per-phase face values / fluxes [9][2N]
};
// planes of a set's float block
__host__ __device__ inline float *pp32_U(float *f, int q, int64_t N) { return f + (int64_t)q * N; }
__host__ __device__ inline float *pp32_V(float *f, int q, int64_t N) { return f + (int64_t)(q + 1) * N; }
__host__ __device__ inline float *pp32_fv(float *f, int q, int64_t N) { return f + (int64_t)(q + 2) * N; }
// the other state set: the one-launch step reads the current set (with its
// halos) and writes this one
struct PP32Set {
    double *rho;
    float *f;
};
struct PP32State {
    PP32Planes P;         // the current set, the shared planes and the per-phase scratch
    PP32Set other;        // the second set
    bool phases = false;  // step with one launch per phase instead of the one-launch tile kernel (D2Q9)
    int cus = 256;        // compute units of the device (tile shape choice)
    int tile = MPD_TILE_AUTO;  // tile shape of the one-launch step: 16x15, 32x16, or chosen per mesh
    double tau_u = 0.0;   // the relaxation time of every cell when the state has one (else 0)
};
// tile shape (MPD_TILE_16X15 or MPD_TILE_32X16) of the FP32-bulk pseudopotential step
int pp32_tile_auto(int nx, int ny, int q, int cus);
// float planes of one state set (the 32-bit range of the step kernel's buffer resources)
int pp32_set_planes(int q);
int pp32_alloc(PP32State &S, void **mem, int64_t *bytes, const Consts &C, int64_t N, int q);
void pp32_from64(const PP32State &S, const DevState &D, hipStream_t st);
// stepped: the state has been stepped since the last conversion from the
// explicit state, so its forces derive from Psi (the face forces from the
// stored face Psi); else they are the uploaded ones in the shared planes
void pp32_to64(const PP32State &S, const DevState &D, const Consts &C, bool stepped, hipStream_t st);
void pp32_step(PP32State &S, const Consts &C, bool stepped, hipStream_t st);

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
    bool pp_stepped = false;  // stepped since the last conversion: the forces derive from Psi
    int tile = MPD_TILE_16X15;   // Allen-Cahn: tile shape of the step kernel (mpd_tile)
    int cus = 256;               // compute units of the device
};

// FP32 planes of the derived state (see fused_kernels.hip); phd is the
// conserved phase field in FP64, phi its FP32 copy for the stencils
struct FusedPlanes32 {
    float *h, *f, *phi, *phl, *ppu, *ppv, *gl;
    double *phd;
};
// one FP32-bulk step (fused_fp32.hip) with tiles of shape `tile` (mpd_tile) on
// a GPU with `cus` compute units
void fused32_step(const FusedPlanes32 &in, const FusedPlanes32 &out, const Consts &C, int64_t N, int tile,
                  int cus, hipStream_t st);
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
