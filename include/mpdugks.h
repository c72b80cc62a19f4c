/*
 * mpdugks.h - C ABI of the MI355X DUGKS time-step library (libmpdugks_hip.so).
 *
 * The library replaces the body of the reference's `DUGKS2DSolver()` time loop
 * (Ark/Cpp/DUGKS2DSolver.cpp:164-328) for a uniform Cartesian mesh: the host
 * driver keeps mesh set-up, initial conditions and output, uploads the state
 * once, advances it on the GPU and downloads it at output cadence.
 *
 * Conventions
 *  - Structured cell index  n = j*nx + i  (i along x, j along y), i.e. the
 *    reference's CarCellArray[i+1][j+1] (Ark/Cpp/Mesh/MeshCartesian.cpp:369-408).
 *  - Populations are planes: tilde[q*ncell + n], q = 0..Q-1 in the reference's
 *    velocity order (D2Q9: Ark/Lib/D2Q9.h:24-26; D2V16: Ark/Lib/D2V16.cpp:54-67).
 *  - Cell macros are planes macro[m*ncell + n], m in mpd_macro_field below; the
 *    fields are the reference's MacroQuantity members (Ark/Header/Macroscopic.h:32-44)
 *    plus the per-cell relaxation time f.tau (Ark/Header/Cell_2D.h:114).
 *  - Every entry point returns 0 on success or a negative mpd_status; nothing
 *    blocks, aborts or throws across the ABI.  mpd_last_error() gives the text.
 *  - A handle is not thread-safe; calls on one handle must come from one thread.
 */
#ifndef MPDUGKS_H
#define MPDUGKS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPDUGKS_ABI_VERSION 1

typedef enum mpd_status {
    MPD_OK = 0,
    MPD_ERR_INVALID = -1,     /* bad argument or unsupported combination */
    MPD_ERR_DEVICE = -2,      /* HIP runtime error */
    MPD_ERR_NONFINITE = -3,   /* NaN/Inf or non-positive density detected */
    MPD_ERR_NOMEM = -4,       /* device allocation failed / over HBM budget */
    MPD_ERR_STATE = -5        /* call not valid in the handle's current state */
} mpd_status;

/* Velocity model: the reference's link-time `QModel` choice (Ark/Makefile:5). */
typedef enum mpd_model {
    MPD_MODEL_AC_P = 0,      /* Ark/Lib/D2Q9AC.cpp  : Allen-Cahn + pressure form, U == 0 */
    MPD_MODEL_AC_HCZ = 1,    /* Ark/Lib/D2Q9HCZ.cpp : Allen-Cahn + HCZ momentum (default) */
    MPD_MODEL_PP_D2Q9 = 2,   /* Ark/Lib/D2Q9PP.cpp  : pseudopotential, Carnahan-Starling */
    MPD_MODEL_PP_D2V16 = 3   /* Ark/Lib/D2V16.cpp   : 16-velocity set with pseudopotential */
} mpd_model;

/* Face reconstruction: Ark/Constant/ZeroFlip.h:27-35.  CD with every model;
 * UW with the Allen-Cahn models (periodic or walled y) and the pseudopotential
 * D2Q9 model (periodic), in both precisions; not with D2V16. */
typedef enum mpd_flux { MPD_FLUX_CD = 0, MPD_FLUX_UW = 1 } mpd_flux;

/* Boundary in y (x is always periodic): Ark/Constant/ZeroFlip.h:59-77. */
typedef enum mpd_bc {
    MPD_BC_PERIODIC = 0,  /* periodic pair 12/8 (GhostCells.cpp:20-28) */
    MPD_BC_WALL_BB = 1,   /* no-slip wall, bounce-back (Wall_3_BB, Boundary.cpp:118-139) */
    MPD_BC_WALL_NEE = 2   /* no-slip wall, non-equilibrium extrapolation (Wall_3_NEE, Boundary.cpp:140-184) */
} mpd_bc;

/* Arithmetic of the bulk update. */
typedef enum mpd_precision {
    MPD_FP64 = 0,            /* plain FP64 drop-in */
    MPD_FP32_BULK = 1        /* FP32 populations/arithmetic, FP64 EOS and global sums: Allen-Cahn
                                models (CD, UW; periodic or walled y), pseudopotential D2Q9 (CD, UW)
                                and D2V16 (CD), periodic */
} mpd_precision;

/* Run-time replacement for the compile-time constants of
 * Ark/Constant/ZeroReference.h and Ark/Constant/ZeroNamespace.h. */
typedef struct mpd_config {
    int32_t nx, ny;
    int32_t model;           /* mpd_model */
    int32_t flux;            /* mpd_flux */
    int32_t bc_y;            /* mpd_bc */
    int32_t precision;       /* mpd_precision */
    double dx;               /* MinL (cell side)                        ZeroReference.h:20 */
    double dt;               /* CFL*MinL/MaxU                           ZeroNamespace.h:10-14 */
    double rt;               /* R0*T0                                   ZeroReference.h:70-74 */
    double phi_l, phi_v;     /* PhiL, PhiV                              ZeroNamespace.h:79 */
    double rho_l, rho_v;     /* RhoL, RhoV                              ZeroNamespace.h:81 */
    double nu_l, nu_v;       /* NuL, NuV (kinematic)                    ZeroNamespace.h:89 */
    double mobility;         /* M_Phi; tau_h = M/RT                     ZeroNamespace.h:63,97 */
    double width;            /* interface width wI                      ZeroNamespace.h:69 */
    double sigma;            /* surface tension                         ZeroNamespace.h:71 */
    double tau0;             /* Tau0 used by the AC-p viscosity         ZeroReference.h:100 */
    double gx, gy, rho_ref;  /* body force g*(rho - rho_ref); 0 in the reference (SourceM.cpp:32-33) */
    double pp_tr;            /* T/Tc of the Carnahan-Starling EOS       ZeroNamespace.h:142 */
    double pp_k;             /* EOS prefactor K                         ZeroNamespace.h:140 */
    double pp_tau;           /* constant relaxation time of the PP model (Preprocess.cpp:376,397) */
} mpd_config;

/* Cell macro planes (order of the `macro` arrays). */
typedef enum mpd_macro_field {
    MPD_M_PHI = 0,   /* Phi */
    MPD_M_RHO,       /* Rho */
    MPD_M_U,         /* U */
    MPD_M_V,         /* V */
    MPD_M_P,         /* p */
    MPD_M_PHIX,      /* Phi_x: grad(phi) inside MacroSource, then the AC source vector */
    MPD_M_PHIY,      /* Phi_y */
    MPD_M_RHOX,      /* Rho_x */
    MPD_M_RHOY,      /* Rho_y */
    MPD_M_FX,        /* Fx */
    MPD_M_FY,        /* Fy */
    MPD_M_LAP,       /* laplacianPhi */
    MPD_M_PPU,       /* prevPhiU */
    MPD_M_PPV,       /* prevPhiV */
    MPD_M_PSI,       /* Psi (pseudopotential) */
    MPD_M_SCG,       /* scG = sign(p - rho*RT) */
    MPD_M_MU,        /* Mu */
    MPD_M_TAU,       /* f.tau of the cell */
    MPD_M_COUNT
} mpd_macro_field;

/* Kernel path used by mpd_step. */
typedef enum mpd_path {
    MPD_PATH_AUTO = 0,       /* fused step kernel where available, else phases */
    MPD_PATH_PHASES = 1,     /* one kernel per reference phase (the plain FP64 path; for an FP32-bulk
                                pseudopotential handle, the FP32 kernels of each phase) */
    MPD_PATH_FUSED = 2       /* one fused kernel per step */
} mpd_path;

/* Global moments returned by mpd_reductions (FP64, compensated sums). */
typedef struct mpd_sums {
    double sum_rho;          /* sum of rho over cells */
    double sum_rho_u;        /* sum of rho*u */
    double sum_rho_v;        /* sum of rho*v */
    double sum_phi;          /* sum of phi (AC models) */
    double kinetic_energy;   /* 0.5 * sum rho*(u^2+v^2) */
    double max_speed;        /* max |u| */
    double residual;         /* sqrt(sum (rho-rho_1k)^2 / (sum rho^2 + 1e-30)); updates rho_1k
                                (Ark/Cpp/DUGKS2DSolver.cpp:807-820) when update_residual != 0 */
    int64_t nonfinite;       /* cells with NaN/Inf or rho <= 0 */
} mpd_sums;

typedef struct mpd_handle mpd_handle;

/* Library ABI version (MPDUGKS_ABI_VERSION). */
int mpd_abi_version(void);

/* Fill `cfg` with the reference's shipped constants for `model`
 * (128x128, dt = 0.5, RT = 1/3, sigma = 1e-2, W = 5, ...). */
int mpd_default_config(int32_t model, mpd_config *cfg);

/* Create a solver on the current HIP device.  Replaces the allocation in
 * AllocateResource (Ark/Cpp/main.cpp:150-186) for the step's state. */
int mpd_create(const mpd_config *cfg, mpd_handle **out);
int mpd_destroy(mpd_handle *h);

/* Select the kernel path (mpd_path); valid before the first mpd_step. */
int mpd_set_path(mpd_handle *h, int32_t path);

/* Visiting order of the pseudopotential face force.  The reference computes
 * each face's Psi and reads its neighbour faces' Psi inside one OpenMP loop
 * (Ark/Cpp/DUGKS2DSolver.cpp:219-230, 733-746; Ark/Lib/GradSchemeBasic.h:181-208),
 * so its result depends on the visiting order; with one thread a neighbour
 * visited earlier contributes its new Psi and one visited later the Psi of the
 * previous step.  MPD_FACE_ORDER_JACOBI (default; every path): all face Psi
 * first, then all forces.  MPD_FACE_ORDER_REFERENCE (FP64 pseudopotential D2Q9,
 * CD or UW faces, periodic; phase path): the reference's one-thread order for a
 * mesh whose faces are numbered as mpdugks_amd.mesh.write_cas writes them
 * (interior x-faces by rows, interior y-faces by rows, then the periodic zones
 * left, right, bottom, top) -- bit for bit the reference's OMP_NUM_THREADS=1 run.
 * With UW faces the two stored copies of each periodic face carry their own Psi
 * and force, which the downloadable state does not hold: after an upload both
 * copies start from the first copy's (exact with CD faces, where they agree). */
typedef enum mpd_face_order { MPD_FACE_ORDER_JACOBI = 0, MPD_FACE_ORDER_REFERENCE = 1 } mpd_face_order;
int mpd_set_face_order(mpd_handle *h, int32_t order);

/* Face Psi of the last step (x-faces [ncell] then y-faces [ncell], layout of
 * the face forces), the state the reference order reads for faces not yet
 * visited (MacroQuantity::Psi of each Face_2D).  mpd_upload_state sets it to
 * 0, as the reference's initialisers leave it.  FP64 pseudopotential D2Q9. */
int mpd_upload_face_psi(mpd_handle *h, const double *psi);
int mpd_download_face_psi(mpd_handle *h, double *psi);

/* Tile shape of the one-launch step kernels (cells per workgroup): the
 * FP32-bulk Allen-Cahn kernel has all three with CD faces and 16x15 with UW
 * faces (a 3-cell halo), the FP32-bulk pseudopotential
 * kernel 16x15 and 32x16 (D2Q9 with CD faces; 16x15 otherwise), the FP64
 * Allen-Cahn kernel 16x15 and 16x16 (other
 * shapes are refused with MPD_ERR_INVALID).
 * MPD_TILE_AUTO (the default) chooses per mesh; the others force one.  The
 * state does not depend on it (bit for bit with periodic and BB walls, to
 * FP32 rounding with NEE walls).  Takes effect at the next mpd_step;
 * mpd_get_tile returns the shape that step uses.  No counterpart in the
 * reference (a launch shape of this library). */
typedef enum mpd_tile {
    MPD_TILE_AUTO = 0,
    MPD_TILE_16X15 = 1,      /* 256 threads */
    MPD_TILE_16X16 = 2,      /* 256 threads */
    MPD_TILE_32X16 = 3       /* 512 threads */
} mpd_tile;
int mpd_set_tile(mpd_handle *h, int32_t tile);
int mpd_get_tile(mpd_handle *h, int32_t *tile);

/* Upload host state (FP64).  tilde_h may be NULL for the PP models;
 * face_fx/face_fy (PP only, may be NULL otherwise) are the persistent face
 * forces of the x-faces (face i,j between cells (i-1,j),(i,j)) followed by the
 * y-faces (face i,j between (i,j-1),(i,j)), 2*ncell each.
 * Mirrors the reference state consumed by the loop: Cell_2D::{h,f}.Tilde and
 * MacroQuantity (Ark/Header/Cell_2D.h:106-149). */
int mpd_upload_state(mpd_handle *h, const double *tilde_h, const double *tilde_f,
                     const double *macro, const double *face_fx, const double *face_fy);

/* Download the state in the same layout (any pointer may be NULL to skip). */
int mpd_download_state(mpd_handle *h, double *tilde_h, double *tilde_f, double *macro,
                       double *face_fx, double *face_fy);

/* Advance nsteps time steps (the `while` body, DUGKS2DSolver.cpp:164-328).
 * Queues the step kernels on the handle's stream and returns without waiting
 * (except with per-launch timing on, and every 256 calls when nothing else
 * synchronises, to read back the timing events); mpd_synchronize,
 * mpd_download_state and mpd_reductions wait for them.  The first call after
 * an upload on an FP64 or Allen-Cahn fused handle also allocates the phase
 * path's scratch. */
int mpd_step(mpd_handle *h, int64_t nsteps);

/* Global sums (SumRho / Residual, DUGKS2DSolver.cpp:793-879).  Synchronises. */
int mpd_reductions(mpd_handle *h, int update_residual, mpd_sums *out);

/* Wait for all work queued on the handle. */
int mpd_synchronize(mpd_handle *h);

/* Bytes of device memory the handle owns. */
int mpd_device_bytes(const mpd_handle *h, int64_t *bytes);

/* Enable (1) or disable (0) per-launch timing: each step's kernels are then
 * bracketed by HIP events on the handle's stream (fused path: exactly the one
 * step kernel).  Off by default. */
int mpd_set_kernel_timing(mpd_handle *h, int enable);

/* Average duration (ms) per step of the step kernels since the last call:
 * with per-launch timing on, the mean of the bracketed launches; otherwise the
 * time on the stream across each mpd_step call's launches divided by the
 * launches counted.  On the fused path a call of two or more steps is timed
 * from the end of its first step kernel (the first kernel's start waits for
 * the host to enqueue it) and counts the others; priming steps after an
 * upload are not counted.  `launches` receives the launches counted.  Resets
 * the counters. */
int mpd_kernel_timing(mpd_handle *h, double *avg_ms, int64_t *launches);

/* ---- batched ensembles ----
 * A batch advances many FP32-bulk cases (Allen-Cahn, CD; same model and y
 * boundary -- periodic, BB or NEE -- any sizes and constants) with one kernel
 * launch per step: every workgroup takes one tile (16x15, 16x16 or 32x16
 * cells, mpd_batch_set_tile) of one case from a work list, and the step pair
 * is captured once in a HIP graph and replayed.  Members keep their own state;
 * mpd_download_state / mpd_reductions on a member synchronise the batch; a
 * member cannot be stepped, uploaded to, re-initialised, switched to another
 * path or given a tile shape on its own (MPD_ERR_STATE).  No counterpart in the reference, which runs
 * one case per process. */
typedef struct mpd_batch mpd_batch;
int mpd_batch_create(mpd_batch **out);
int mpd_batch_destroy(mpd_batch *b);
/* h must have taken >= 2 steps on the fused path (MPD_ERR_STATE otherwise). */
int mpd_batch_add(mpd_batch *b, mpd_handle *h);
int mpd_batch_remove(mpd_batch *b, mpd_handle *h);
int mpd_batch_size(const mpd_batch *b, int32_t *ncases, int64_t *ncells);
/* Advance every member by nsteps (synchronises at the end). */
int mpd_batch_step(mpd_batch *b, int64_t nsteps);
/* Mean wall time per batch step on the batch stream since the last call. */
int mpd_batch_timing(mpd_batch *b, double *avg_ms, int64_t *steps);
/* Tile shape of the batched launch (mpd_tile).  MPD_TILE_AUTO (the default)
 * tunes it: the first mpd_batch_step of at least 24 steps after the members
 * change runs 8 steps with each of the three shapes (ordinary steps of the
 * run: with periodic and BB walls all give the same state bit for bit), times
 * the last 6 of each and keeps the fastest.  Any other value forces a shape.
 * Batches with NEE walls use 16x15 unless forced, so that their results do
 * not depend on a timing.  mpd_batch_tile returns the shape in use and the
 * last tuning's ms per step of each shape, indexed by mpd_tile - 1 (0 where
 * not measured).  No counterpart in the reference. */
int mpd_batch_set_tile(mpd_batch *b, int32_t tile);
int mpd_batch_tile(const mpd_batch *b, int32_t *tile, double ms_per_step[3]);

/* ---- host-side initial states (upload layout; no device work) ---- */

/* AC_Drop (Ark/Cpp/Preprocess.cpp:705-760, AnalyticalPhiAC_Drop :402-419) on the
 * mesh of `cfg`: phi = (PhiL+PhiV)/2 + sign*(PhiL-PhiV)/2*tanh(2(r-R)/W).
 * sign = +1 is the reference's case (light phase inside); -1 gives a droplet.
 * Macros start from p = p0, U = V = 0, Mu = mu0, f.tau = mu0/(rho RT). */
int mpd_init_ac_drop(const mpd_config *cfg, double cx, double cy, double radius, int32_t sign,
                     double p0, double mu0, double *tilde_h, double *tilde_f, double *macro);

/* Interface along x at per-column heights y0[nx] (no counterpart in the
 * reference; used for the Rayleigh-Taylor case): phi = tanh profile of width W,
 * sign = +1 puts phase L above; Mu from the HCZ viscosity law. */
int mpd_init_ac_interface(const mpd_config *cfg, const double *y0, int32_t sign, double p0,
                          double *tilde_h, double *tilde_f, double *macro);

/* SCMP_Drop (Ark/Cpp/Preprocess.cpp:337-401) for the pseudopotential model. */
int mpd_init_pp_drop(const mpd_config *cfg, double cx, double cy, double radius, double rho_l,
                     double rho_v, double width, double mu0, double *tilde_f, double *macro,
                     double *face_fx, double *face_fy);

/* SCMP_FlatInterface (Ark/Cpp/Preprocess.cpp:291-336): a liquid band,
 * rho = rho_v + (rho_l - rho_v)/2 (tanh(slope (y - y_lo)) - tanh(slope (y - y_hi))),
 * u = 0, tilde = equilibrium; the reference's case is slope = 0.4,
 * y_lo = ChLength/4, y_hi = 3 ChLength/4. */
int mpd_init_pp_flat(const mpd_config *cfg, double y_lo, double y_hi, double slope, double rho_l,
                     double rho_v, double mu0, double *tilde_f, double *macro, double *face_fx,
                     double *face_fy);

/* Pseudopotential liquid-vapour start: rho = rho0*(1 + U(-amp, amp)) per cell,
 * std::mt19937_64(seed), u = 0, tilde = equilibrium. */
int mpd_init_pp_noise(const mpd_config *cfg, double rho0, double amp, uint64_t seed, double mu0,
                      double *tilde_f, double *macro, double *face_fx, double *face_fy);

/* ---------------- start states built on the device ----------------
 * The handle's state is computed on its GPU (no host arrays, no upload), as the
 * host initialisers below compute it, with the device's tanh/sqrt (agreement to
 * a few units in the last place).  Afterwards the handle is as after
 * mpd_upload_state.  For drivers that create many cases (Preprocess.cpp:705-823
 * builds them on the host cell by cell). */

/* AC_Drop for 1..8 drops (Ark/Cpp/Preprocess.cpp:705-760): each drop's tanh
 * profile of mpd_init_ac_drop, joined as the larger phi for droplets
 * (sign = -1) and the smaller for bubbles (sign = +1). */
int mpd_device_init_drops(mpd_handle *h, int32_t ndrops, const double *cx, const double *cy, const double *radius,
                          int32_t sign, double p0, double mu0);

/* mpd_init_ac_interface on the device; y0[nx] is a host array. */
int mpd_device_init_interface(mpd_handle *h, const double *y0, int32_t sign, double p0);

/* mpd_device_init_interface without any device allocation: the y0 staging
 * buffer is a fixed device array of the library (16384 columns; larger nx is
 * refused with MPD_ERR_INVALID) instead of a per-call hipMalloc/hipFree.  Same
 * state bit for bit.  Calls that overlap on one GPU share that array and race.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (serving loops that must not allocate per request, code running
 * while allocations are forbidden). */
int mpd_device_init_interface_noalloc(mpd_handle *h, const double *y0, int32_t sign, double p0);

/* Pseudopotential noisy start: rho = rho0 (1 + U(-amp, amp)) per cell with a
 * counter-based splitmix64 stream of (seed, cell) (not the host's mt19937_64
 * sequence), u = 0, tilde = equilibrium, face forces and face Psi 0. */
int mpd_device_init_pp_noise(mpd_handle *h, double rho0, double amp, uint64_t seed, double mu0);

/* mpd_init_pp_drop (SCMP_Drop) and mpd_init_pp_flat (SCMP_FlatInterface) on the
 * device; face forces and the previous step's face Psi start at 0. */
int mpd_device_init_pp_drop(mpd_handle *h, double cx, double cy, double radius, double rho_l, double rho_v,
                            double width, double mu0);
int mpd_device_init_pp_flat(mpd_handle *h, double y_lo, double y_hi, double slope, double rho_l, double rho_v,
                            double mu0);

/* ---------------- output files of the reference driver ----------------
 * Byte-identical writers of the files the reference's driver writes
 * (Ark/Cpp/Output.cpp), fed from downloaded macro planes.  Host code: no GPU
 * is needed.  Paths are used as given; mpd_output_name gives the reference's
 * file name relative to its FlowField directory. */
typedef enum mpd_output_kind {
    MPD_OUT_FLOWFIELD = 0,  /* global/step..   Output_Flowfield, Output.cpp:484-555 */
    MPD_OUT_SUMRHO = 1,     /* Convergence/SumRho_..   Output_SumRho, :258-271 */
    MPD_OUT_RESIDUAL = 2,   /* Convergence/Residual_.. Output_Residual, :272-287 */
    MPD_OUT_MIDX = 3,       /* Convergence/MidX_..     Output_MidX, :309-390 */
    MPD_OUT_MIDY = 4,       /* Convergence/MidY_..     Output_MidY, :404-458 */
    MPD_OUT_DIAG = 5        /* Convergence/Diag_..     Output_Diag, :459-515 */
} mpd_output_kind;

/* Names and constants the writers print (the reference's compile-time values). */
typedef struct mpd_output_info {
    double ma, cfl, mu0, re;  /* Ma, CFL, Mu0, Re (ZeroReference.h, ZeroNamespace.h) */
    double m_phi;             /* PhaseFieldAC::M_Phi, printed in profile zone titles */
    double rho_l, rho_v;      /* density ratio in profile file names */
    double temperature;       /* the T zone of the flow field (MacroQuantity::T = T0) */
    int32_t mesh_nl;          /* NL, the mesh-size tag of file names (not `nl`: the reference
                                 defines `nl` as a macro, Ark/Header/ZeroMacros.h:19) */
    char mesh_type[32];       /* _MESHTYPE_ARK */
    char mesh_file[128];      /* _MESHFILE_NAME_ARK */
} mpd_output_info;

/* The reference's defaults for `cfg` (ZeroFlip.h, ZeroReference.h). */
int mpd_default_output_info(const mpd_config *cfg, mpd_output_info *info);

/* File name of `kind` for `step` at time `t` (NUL-terminated, at most `len` bytes). */
int mpd_output_name(int32_t kind, const mpd_output_info *info, int64_t step, double t, char *buf,
                    int32_t len);

/* Output_Flowfield: zones Rho, U V, p, T, Phi, f.tau, Fx, Fy over the cells in
 * the driver's CellArray order (`order[k]` = structured index of cell k; NULL for
 * structured order), scientific with 12 digits. */
int mpd_write_flowfield(const char *path, int32_t nx, int32_t ny, const double *macro, const int64_t *order,
                        const mpd_output_info *info);

/* Output_MidX / Output_MidY / Output_Diag (kind MPD_OUT_MIDX/MIDY/DIAG): Rho, p,
 * thermodynamic pressure (thermoPressure, Output.cpp:146-162), U, V along the
 * profile; cfg gives the mesh and the phase-field constants. */
int mpd_write_profile(const char *path, int32_t kind, const mpd_config *cfg, const double *macro,
                      const mpd_output_info *info);

/* Output_SumRho / Output_Residual: append one "t    value" record. */
int mpd_append_convergence(const char *path, double t, double value);

/* Case.ark (OutputCase, Output.cpp:65-144). */
typedef struct mpd_case_info {
    char case_name[64], mesh_file[128], mesh_type[32], flux_scheme[16], bc[32];
    char velocity_model[16], force_model[32], eos_model[32];
    int32_t zone_left, zone_right, zone_top, zone_bottom;
    double omega0, prandtl, nk, cv, gamma;
    int32_t qu, qv, mesh_nl;  /* Qu, Qv, NL */
    double ch_length, min_l, kn, ma, re, u0, r0, t0, mu0, nu0, tau0, lambda0, cfl, dt, residual;
    int32_t end_step;
    double maspan, eta;
    int32_t phase_field;      /* print the phase-field block */
    double cn, pe, eo, mo, re_mp, mobility, tau_mass, width, sigma, beta, kappa;
    double rho_l, rho_v, mu_l, mu_v, nu_l, nu_v, radius;
} mpd_case_info;

/* The reference's Case.ark values for `cfg` (AC_Drop case, shipped flags). */
int mpd_default_case_info(const mpd_config *cfg, mpd_case_info *info);
int mpd_write_case(const char *path, const mpd_case_info *info);

/* Text of the last error on this thread. */
const char *mpd_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* MPDUGKS_H */
