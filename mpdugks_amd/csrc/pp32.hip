// FP32-bulk pseudopotential D2Q9 step (periodic mesh, CD faces).
//
// The phases of the reference step for the D2Q9PP model, one kernel each as in
// the FP64 phase path (phase_kernels.hip: k_phase1, k_pp_face_psi,
// k_pp_face_flux, k_update, k_ppforce), with populations, bar-plus, face
// values, face macros and fluxes in FP32.  The Carnahan-Starling pressure and
// Psi = sqrt(2 (p - rho RT) / scG) are evaluated in FP64 from the FP32
// density: near the spinodal p - rho RT is a difference of nearly equal
// numbers.  The global sums stay FP64 (mpd_reductions on the materialised
// state).
//
// State (structure of arrays, cell n = j*nx + i; face t < N is the x-face west
// of cell t, face N + t the y-face south of cell t):
//   f~[9][N], bar-plus[9][N]
//   cell rho (FP64), U, V, tau, Fx, Fy, Psi, scG [N]  (the reference's MacroQuantity)
//   face force Fx, Fy [2N]                      (persistent per face, D2Q9PP.cpp:117-118)
//   face Psi, rho, U, V, scG, mass flux [2N], face fluxes [9][2N]  (scratch of one step)
//
// Mass.  The density is the conserved variable, kept in FP64: it moves only by
// the face mass fluxes (each face's sum over populations, shared by both cells
// with opposite signs) and f~_0 takes the remainder, so that sum_q f~ = rho.
// In exact arithmetic this is the reference's update; in FP32 it keeps the
// rounding of the per-population products (and of the FP32 weights, whose sum
// is not exactly 1) from drifting the total mass.
#include <hip/hip_runtime.h>

#include "fused.h"
#include "physics.h"

namespace mpd {

namespace {

__device__ __forceinline__ int wrapp(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

inline unsigned blocks(size_t n) { return (unsigned)((n + 255) / 256); }

// Update_PseudoForce(cell) (PseudoP.cpp:31-38): F = -scG Psi grad(Psi), from
// the Psi that the last update left
__device__ __forceinline__ void pp32_cell_force(const PP32Planes &P, const Consts &C, size_t n, float &Fx, float &Fy)
{
    const int nx = P.nx, ny = P.ny;
    const int i = (int)(n % nx), j = (int)(n / nx);
    auto S = [&](int a, int b) { return P.psi[(size_t)wrapp(b, ny) * nx + wrapp(a, nx)]; };
    const float c[4] = {S(i + 1, j), S(i, j + 1), S(i - 1, j), S(i, j - 1)};
    const float d[4] = {S(i + 1, j + 1), S(i - 1, j + 1), S(i - 1, j - 1), S(i + 1, j - 1)};
    float gx, gy;
    grad9<float>(C, c, d, gx, gy);
    const float sp = -P.scg[n] * P.psi[n];
    Fx = gx * sp;
    Fy = gy * sp;
}

// phase 1: equilibria, He-Shan-Doolen source and bar-plus of every cell
// (Update_phi_Eq / Update_DVDF_Source / Update_phi_BP: D2Q9PP.cpp:51-62,137-145,
// DUGKS2DSolver.cpp:379-401).  FORCE: first the cell force that the previous
// step's last phase (Update_PseudoForce) produced, from Psi, and store it for
// the update; without FORCE the stored force is used (the first step after a
// conversion from the explicit state, whose force may not derive from Psi).
template <bool FORCE>
__global__ void __launch_bounds__(256) k_pp32_phase1(PP32Planes P, Consts C)
{
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t N = P.N;
    if (n >= N) return;
    CellMacro<float> cm = {};
    cm.rho = (float)P.rho[n];
    cm.U = P.U[n];
    cm.V = P.V[n];
    cm.tau = P.tau[n];
    if (FORCE) {
        pp32_cell_force(P, C, n, cm.Fx, cm.Fy);
        P.Fx[n] = cm.Fx;
        P.Fy[n] = cm.Fy;
    } else {
        cm.Fx = P.Fx[n];
        cm.Fy = P.Fy[n];
    }
    float ft[Q9], fb[Q9], ht[Q9] = {}, hb[Q9];
#pragma unroll
    for (int q = 0; q < Q9; ++q) ft[q] = P.f[q * N + n];
    cell_barplus<float, MPD_MODEL_PP_D2Q9>(C, cm, ht, ft, hb, fb);
#pragma unroll
    for (int q = 0; q < Q9; ++q) P.fbp[q * N + n] = fb[q];
}

// CD face value of population q (CD_Interior_phi_Bh, DUGKS2DSolver.cpp:604-687)
// on face t of direction dir at (i, j): owner west / south, neighbour east / north
__device__ __forceinline__ float pp32_face_value(const PP32Planes &P, const Consts &C, int dir, int q, int i, int j)
{
    const int nx = P.nx, ny = P.ny;
    const float *b = P.fbp + (size_t)q * P.N;
    auto B = [&](int a, int c) { return b[(size_t)wrapp(c, ny) * nx + wrapp(a, nx)]; };
    if (dir == 0) {
        const int oi = i - 1;
        return cd_face_value<float>(C, 0, q, B(oi, j), B(i, j), B(i, j + 1), B(i, j - 1), B(oi, j + 1), B(oi, j - 1));
    }
    const int oj = j - 1;
    return cd_face_value<float>(C, 1, q, B(i, oj), B(i, j), B(i + 1, j), B(i - 1, j), B(i + 1, oj), B(i - 1, oj));
}

// faces, pass 1: face macros (Update_MacroVar_h, D2Q9PP.cpp:110-136) with the
// face force of the previous step, and Psi of the face (Update_PseudoPsi,
// PseudoP.cpp:23-30) in FP64
__global__ void __launch_bounds__(256) k_pp32_face_psi(PP32Planes P, Consts C)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t N = P.N;
    if (t >= 2 * N) return;
    const int dir = t < N ? 0 : 1;
    const size_t fi = dir ? t - N : t;
    const int i = (int)(fi % P.nx), j = (int)(fi / P.nx);
    float fb[Q9];
    float *fv = P.ffl + (dir ? (size_t)Q9 * N : 0);
#pragma unroll
    for (int q = 0; q < Q9; ++q) {
        fb[q] = pp32_face_value(P, C, dir, q, i, j);
        fv[(size_t)q * N + fi] = fb[q];  // the flux pass reads it back and overwrites it with the flux
    }
    float s = 0.f, sx = 0.f, sy = 0.f;
#pragma unroll
    for (int q = 0; q < Q9; ++q) s += fb[q];
#pragma unroll
    for (int q = 0; q < Q9; ++q) sx += fb[q] * (float)C.xi[0][q];
#pragma unroll
    for (int q = 0; q < Q9; ++q) sy += fb[q] * (float)C.xi[1][q];
    const float rho = s;
    float U = sx / rho, V = sy / rho;
    U += 0.5f * (float)C.hdt * P.ffx[t] / rho;
    V += 0.5f * (float)C.hdt * P.ffy[t] / rho;
    const double p = C.pp_k * carnahan_starling((double)rho, C.t_cs);
    double scg, psi;
    calc_psi(p, (double)rho, C.rt, scg, psi);
    P.fpsi[t] = (float)psi;
    P.frho[t] = rho;
    P.fU[t] = U;
    P.fV[t] = V;
    P.fscg[t] = (float)scg;
}

// faces, pass 2: face force from the face-lattice gradient of Psi
// (update_MQ_x/y(Face_2D*), GradSchemeBasic.h:181-208; Update_PseudoForce(face),
// PseudoP.cpp:39-46), half-step BGK (DUGKS2DSolver.cpp:402-423) and micro-flux
// (:424-441) of the face values that pass 1 left in the flux planes
__global__ void __launch_bounds__(256) k_pp32_face_flux(PP32Planes P, Consts C)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t N = P.N;
    if (t >= 2 * N) return;
    const int dir = t < N ? 0 : 1;
    const size_t fi = dir ? t - N : t;
    const int nx = P.nx, ny = P.ny;
    const int i = (int)(fi % nx), j = (int)(fi / nx);
    const float *ps = P.fpsi + (dir ? N : 0);
    auto S = [&](int a, int b) { return ps[(size_t)wrapp(b, ny) * nx + wrapp(a, nx)]; };
    const float c[4] = {S(i + 1, j), S(i, j + 1), S(i - 1, j), S(i, j - 1)};
    const float d[4] = {S(i + 1, j + 1), S(i - 1, j + 1), S(i - 1, j - 1), S(i + 1, j - 1)};
    float gx, gy;
    grad9<float>(C, c, d, gx, gy);
    const float sp = -P.fscg[t] * P.fpsi[t];
    gx *= sp;
    gy *= sp;
    P.ffx[t] = gx;
    P.ffy[t] = gy;
    FaceMacro<float> fm = {};
    fm.rho = P.frho[t];
    fm.U = P.fU[t];
    fm.V = P.fV[t];
    fm.Fx = gx;
    fm.Fy = gy;
    const float ah = (float)C.ah_pp, bh = (float)C.bh_pp, ch = (float)C.ch_pp;
    float *fl = P.ffl + (dir ? (size_t)Q9 * N : 0);
    float mass = 0.f;
#pragma unroll
    for (int q = 0; q < Q9; ++q) {
        const float fb = fl[(size_t)q * N + fi];
        float hh, ff;
        face_bgk<float, MPD_MODEL_PP_D2Q9>(C, fm, q, 0.f, fb, ah, bh, ch, hh, ff);
        const float xin = (float)(C.dx * C.xi[dir][q]);
        fl[(size_t)q * N + fi] = ff * xin;
        mass += ff * xin;
    }
    P.fmass[t] = mass;
}

// Update_phi_T (DUGKS2DSolver.cpp:748-792), Update_MacroVar (D2Q9PP.cpp:89-109)
// and Psi of the cell (FP64 EOS)
__global__ void __launch_bounds__(256) k_pp32_update(PP32Planes P, Consts C)
{
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t N = P.N;
    if (n >= N) return;
    const int nx = P.nx, ny = P.ny;
    const int i = (int)(n % nx), j = (int)(n / nx);
    const size_t kE = (size_t)j * nx + wrapp(i + 1, nx), kW = n;
    const size_t kN = (size_t)wrapp(j + 1, ny) * nx + i, kS = n;
    const float aTP = 4.f / 3.f, bTP = 1.f / 3.f, dtv = (float)C.dtv;
    const float *fx = P.ffl, *fy = P.ffl + (size_t)Q9 * N;
    float ft[Q9];
    double rest = 0.0;
#pragma unroll
    for (int q = 0; q < Q9; ++q) {
        float s = 0.f;
        s += -fx[q * N + kE];
        s += -fy[q * N + kN];
        s += fx[q * N + kW];
        s += fy[q * N + kS];
        ft[q] = aTP * P.fbp[q * N + n] - bTP * P.f[q * N + n] + dtv * s;
        if (q > 0) rest += (double)ft[q];
    }
    // the conserved density: antisymmetric face mass fluxes, in FP64
    const float *mx = P.fmass, *my = P.fmass + N;
    const double dm = ((double)mx[kW] - (double)mx[kE]) + ((double)my[kS] - (double)my[kN]);
    const double rho_d = P.rho[n] + C.dtv * dm;
    ft[0] = (float)(rho_d - rest);
#pragma unroll
    for (int q = 0; q < Q9; ++q) P.f[q * N + n] = ft[q];
    float sx = 0.f, sy = 0.f;
#pragma unroll
    for (int q = 0; q < Q9; ++q) sx += ft[q] * (float)C.xi[0][q];
#pragma unroll
    for (int q = 0; q < Q9; ++q) sy += ft[q] * (float)C.xi[1][q];
    const float rho = (float)rho_d;
    float U = sx / rho, V = sy / rho;
    U += (float)C.hdt * P.Fx[n] / rho;
    V += (float)C.hdt * P.Fy[n] / rho;
    const double p = C.pp_k * carnahan_starling(rho_d, C.t_cs);
    double scg, psi;
    calc_psi(p, rho_d, C.rt, scg, psi);
    P.rho[n] = rho_d;
    P.U[n] = U;
    P.V[n] = V;
    P.psi[n] = (float)psi;
    P.scg[n] = (float)scg;
}

// the cell force of the current state (the last phase of the reference step),
// for downloads: the steps themselves compute it at the start of the next step
__global__ void __launch_bounds__(256) k_pp32_force(PP32Planes P, Consts C)
{
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (size_t)P.N) return;
    float Fx, Fy;
    pp32_cell_force(P, C, n, Fx, Fy);
    P.Fx[n] = Fx;
    P.Fy[n] = Fy;
}

// explicit FP64 state -> FP32 planes
__global__ void __launch_bounds__(256) k_pp32_from64(PP32Planes P, DevState S)
{
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t N = P.N;
    if (n >= N) return;
    const double *m = S.m;
#pragma unroll
    for (int q = 0; q < Q9; ++q) P.f[q * N + n] = (float)S.f[q * N + n];
    P.rho[n] = m[MPD_M_RHO * N + n];
    P.U[n] = (float)m[MPD_M_U * N + n];
    P.V[n] = (float)m[MPD_M_V * N + n];
    P.tau[n] = (float)m[MPD_M_TAU * N + n];
    P.Fx[n] = (float)m[MPD_M_FX * N + n];
    P.Fy[n] = (float)m[MPD_M_FY * N + n];
    P.psi[n] = (float)m[MPD_M_PSI * N + n];
    P.scg[n] = (float)m[MPD_M_SCG * N + n];
    for (size_t t = n; t < 2 * N; t += N) {
        P.ffx[t] = (float)S.ffx[t];
        P.ffy[t] = (float)S.ffy[t];
    }
}

// FP32 planes -> explicit FP64 state (p from the FP64 EOS of the stored density)
__global__ void __launch_bounds__(256) k_pp32_to64(PP32Planes P, DevState S, Consts C)
{
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t N = P.N;
    if (n >= N) return;
    double *m = S.m;
#pragma unroll
    for (int q = 0; q < Q9; ++q) S.f[q * N + n] = (double)P.f[q * N + n];
    const double rho = P.rho[n];
    m[MPD_M_RHO * N + n] = rho;
    m[MPD_M_U * N + n] = (double)P.U[n];
    m[MPD_M_V * N + n] = (double)P.V[n];
    m[MPD_M_P * N + n] = C.pp_k * carnahan_starling(rho, C.t_cs);
    m[MPD_M_FX * N + n] = (double)P.Fx[n];
    m[MPD_M_FY * N + n] = (double)P.Fy[n];
    m[MPD_M_PSI * N + n] = (double)P.psi[n];
    m[MPD_M_SCG * N + n] = (double)P.scg[n];
    m[MPD_M_TAU * N + n] = (double)P.tau[n];
    for (size_t t = n; t < 2 * N; t += N) {
        S.ffx[t] = (double)P.ffx[t];
        S.ffy[t] = (double)P.ffy[t];
    }
}

}  // namespace

int pp32_alloc(PP32Planes &P, void **mem, int64_t *bytes, const Consts &C, int64_t N)
{
    // FP64 density [N]; FP32 cell planes: f 9, bar-plus 9, U V tau Fx Fy Psi
    // scG 7; FP32 face planes (2N each): force 2, Psi rho U V scG mass 6, fluxes 9
    const size_t nf = (size_t)(9 + 9 + 7) * N + (size_t)(2 + 6 + 9) * 2 * N;
    const size_t bytes_total = (size_t)N * sizeof(double) + nf * sizeof(float);
    void *base = nullptr;
    if (hipMalloc(&base, bytes_total) != hipSuccess) return -1;
    *mem = base;
    *bytes = (int64_t)bytes_total;
    P.nx = C.nx;
    P.ny = C.ny;
    P.N = N;
    const size_t n = (size_t)N;
    P.rho = (double *)base;
    float *p = (float *)(P.rho + n);
    P.f = p; p += 9 * n;
    P.fbp = p; p += 9 * n;
    P.U = p; p += n;
    P.V = p; p += n;
    P.tau = p; p += n;
    P.Fx = p; p += n;
    P.Fy = p; p += n;
    P.psi = p; p += n;
    P.scg = p; p += n;
    P.ffx = p; p += 2 * n;
    P.ffy = p; p += 2 * n;
    P.fpsi = p; p += 2 * n;
    P.frho = p; p += 2 * n;
    P.fU = p; p += 2 * n;
    P.fV = p; p += 2 * n;
    P.fscg = p; p += 2 * n;
    P.fmass = p; p += 2 * n;
    P.ffl = p;
    return 0;
}

void pp32_from64(const PP32Planes &P, const DevState &S, hipStream_t st)
{
    hipLaunchKernelGGL(k_pp32_from64, dim3(blocks(P.N)), dim3(256), 0, st, P, S);
}

void pp32_to64(const PP32Planes &P, const DevState &S, const Consts &C, bool force_stale, hipStream_t st)
{
    if (force_stale) hipLaunchKernelGGL(k_pp32_force, dim3(blocks(P.N)), dim3(256), 0, st, P, C);
    hipLaunchKernelGGL(k_pp32_to64, dim3(blocks(P.N)), dim3(256), 0, st, P, S, C);
}

void pp32_step(const PP32Planes &P, const Consts &C, bool force_stale, hipStream_t st)
{
    const size_t N = P.N;
    if (force_stale) hipLaunchKernelGGL(k_pp32_phase1<true>, dim3(blocks(N)), dim3(256), 0, st, P, C);
    else hipLaunchKernelGGL(k_pp32_phase1<false>, dim3(blocks(N)), dim3(256), 0, st, P, C);
    hipLaunchKernelGGL(k_pp32_face_psi, dim3(blocks(2 * N)), dim3(256), 0, st, P, C);
    hipLaunchKernelGGL(k_pp32_face_flux, dim3(blocks(2 * N)), dim3(256), 0, st, P, C);
    hipLaunchKernelGGL(k_pp32_update, dim3(blocks(N)), dim3(256), 0, st, P, C);
}

}  // namespace mpd
