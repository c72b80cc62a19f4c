// FP32-bulk pseudopotential D2Q9 step (periodic mesh; CD faces, and UW faces on
// the one-launch tile kernel).
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

// FP32 constants of the step, derived on the host in FP64 from Consts: the
// reference's expressions with every division by a constant replaced by a
// product with its reciprocal, and the D2Q9 velocities folded at compile time
// (xi = MaSpan e, e in {-1, 0, 1}^2).
struct KPP {
    float ms;          // MaSpan
    float ms_rt;       // MaSpan / RT
    float inv_rt, lambda0, hdt, dt, half_hdt;
    float ra, rb;      // CD face value, e_n = +1: ra o + rb n, ra = 1/2 + A, rb = 1/2 - A, A = hdt MaSpan / dx
    float rt4;         // tangential factor hdt MaSpan / (2 twodx)
    float inv_d2x6;    // 9-point gradient
    float w[3];        // weights: rest, axis, diagonal
    float ah, bh, ch;  // face BGK factors
    float dtv;         // dt / V
    float flux;        // dx MaSpan: micro-flux per unit e_n
    // UW face value (x_f - h xi - x_c) . grad b / dx with the gradient as 6-point
    // sums (x 12 dx): normal (1/2 - A)/12 upwind, tangential A/12, 1/24 each side for e_n = 0
    float uw_n, uw_t, uw_0;
};

KPP make_kpp(const Consts &C)
{
    KPP k;
    const double ms = C.xi[0][1];
    k.ms = (float)ms;
    k.ms_rt = (float)(ms / C.rt);
    k.inv_rt = (float)(1.0 / C.rt);
    k.lambda0 = (float)C.lambda0;
    k.hdt = (float)C.hdt;
    k.dt = (float)C.dt;
    k.half_hdt = (float)(0.5 * C.hdt);
    const double A = C.hdt * ms / C.dx;
    k.ra = (float)(0.5 + A);
    k.rb = (float)(0.5 - A);
    k.rt4 = (float)(C.hdt * ms / (2.0 * C.twodx));
    k.inv_d2x6 = (float)(1.0 / C.d2x6);
    k.w[0] = (float)C.w[0];
    k.w[1] = (float)C.w[1];
    k.w[2] = (float)C.w[2];
    k.ah = (float)C.ah_pp;
    k.bh = (float)C.bh_pp;
    k.ch = (float)C.ch_pp;
    k.dtv = (float)C.dtv;
    k.flux = (float)(C.dx * ms);
    k.uw_n = (float)((0.5 - A) / 12.0);
    k.uw_t = (float)(A / 12.0);
    k.uw_0 = (float)(1.0 / 24.0);
    return k;
}

// weight class of population q: 0 rest, 1 axis (odd q), 2 diagonal (even q)
__host__ __device__ constexpr int wcls(int q) { return q == 0 ? 0 : ((q & 1) ? 1 : 2); }

// 9-point isotropic gradient (update_MQ_x/y, Ark/Lib/GradSchemeBasic.h:153-180);
// c = E, N, W, S, d = NE, NW, SW, SE
__device__ __forceinline__ void pp_grad9(const KPP &k, const float *c, const float *d, float &gx, float &gy)
{
    gx = (4.f * (c[0] - c[2]) + (d[0] - d[1] + d[3] - d[2])) * k.inv_d2x6;
    gy = (4.f * (c[1] - c[3]) + (d[0] - d[3] + d[1] - d[2])) * k.inv_d2x6;
}

// CD face value of population q (CD_Interior_phi_Bh, DUGKS2DSolver.cpp:604-687)
// on a face of direction DIR: owner o, neighbour n, the neighbour's and the
// owner's tangential pairs (high, low):
//   1/2 (o + n) - hdt xi_n (n - o)/dx - hdt xi_t (1/2 (tn_hi - tn_lo) + 1/2 (to_hi - to_lo))/twodx
template <int DIR>
__device__ __forceinline__ float pp_recon(const KPP &k, int q, float bo, float bn, float tn_hi, float tn_lo,
                                          float to_hi, float to_lo)
{
    const int en = DIR == 0 ? ex9(q) : ey9(q), et = DIR == 0 ? ey9(q) : ex9(q);
    float v = en > 0 ? fmaf(k.rb, bn, k.ra * bo) : (en < 0 ? fmaf(k.ra, bn, k.rb * bo) : 0.5f * (bo + bn));
    if (et != 0) v = fmaf(et > 0 ? -k.rt4 : k.rt4, (tn_hi - tn_lo) + (to_hi - to_lo), v);
    return v;
}

// Equilibrium and He-Shan-Doolen source of population q (Update_phi_Eq /
// Update_DVDF_Source, D2Q9PP.cpp:51-62,137-145): f_eq = w rho (1 + u1 + u1^2/2
// - |u|^2 Lambda0), u1 = xi.u/RT; f_So = F.(xi - u) f_eq / (rho RT).  uul =
// |u|^2 Lambda0, irr = 1/(rho RT).
__device__ __forceinline__ float pp_feq(const KPP &k, int q, float rho, float U, float V, float uul, float Fx,
                                        float Fy, float irr, float &fso)
{
    const int ex = ex9(q), ey = ey9(q);
    const float u1 = k.ms_rt * ((float)ex * U + (float)ey * V);
    const float feq = k.w[wcls(q)] * rho * (1.f + u1 + 0.5f * u1 * u1 - uul);
    fso = (Fx * ((float)ex * k.ms - U) + Fy * ((float)ey * k.ms - V)) * feq * irr;
    return feq;
}

// bar-plus of one cell (Update_phi_BP, DUGKS2DSolver.cpp:379-401; factors
// Cell_2D.cpp:120-125): aBP f~ + bBP f_eq + cBP f_So
__device__ __forceinline__ void pp_barplus(const KPP &k, float rho, float U, float V, float tau, float Fx, float Fy,
                                           const float *ft, float *fb)
{
    const float abp = (2.f * tau - k.hdt) / (2.f * tau + k.dt);
    const float bbp = 1.f - abp, cbp = tau * bbp;
    const float uul = (U * U + V * V) * k.lambda0;
    const float irr = k.inv_rt / rho;
#pragma unroll
    for (int q = 0; q < Q9; ++q) {
        float fso;
        const float feq = pp_feq(k, q, rho, U, V, uul, Fx, Fy, irr, fso);
        fb[q] = abp * ft[q] + bbp * feq + cbp * fso;
    }
}

// Psi and scG of a density (MacroQuantity::calcPsi, Ark/Header/Macroscopic.h:64-68)
// in FP64 (near the spinodal p - rho RT is a difference of nearly equal
// numbers): Psi = sqrt(2 (p - rho RT) / scG) = sqrt(|2 (p - rho RT)|) for
// scG = +-1, NaN for scG = 0 (0/0); rounded to FP32 once.
__device__ __forceinline__ void pp_psi(const Consts &C, double rho, float &scg, float &psi)
{
    const double p = C.pp_k * carnahan_starling(rho, C.t_cs);
    const double rrt = rho * C.rt;
    const int s = (p > rrt) - (p < rrt);
    scg = (float)s;
    psi = s != 0 ? (float)sqrt(fabs(2.0 * (p - rrt))) : __builtin_nanf("");
}

// Update_PseudoForce(cell) (PseudoP.cpp:31-38): F = -scG Psi grad(Psi), from
// the Psi that the last update left
__device__ __forceinline__ void pp32_cell_force(const PP32Planes &P, const KPP &k, size_t n, float &Fx, float &Fy)
{
    const int nx = P.nx, ny = P.ny;
    const int i = (int)(n % nx), j = (int)(n / nx);
    auto S = [&](int a, int b) { return P.psi[(size_t)wrapp(b, ny) * nx + wrapp(a, nx)]; };
    const float c[4] = {S(i + 1, j), S(i, j + 1), S(i - 1, j), S(i, j - 1)};
    const float d[4] = {S(i + 1, j + 1), S(i - 1, j + 1), S(i - 1, j - 1), S(i + 1, j - 1)};
    float gx, gy;
    pp_grad9(k, c, d, gx, gy);
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
__global__ void __launch_bounds__(256) k_pp32_phase1(PP32Planes P, KPP k)
{
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t N = P.N;
    if (n >= N) return;
    float Fx, Fy;
    if (FORCE) {
        pp32_cell_force(P, k, n, Fx, Fy);
        P.Fx[n] = Fx;
        P.Fy[n] = Fy;
    } else {
        Fx = P.Fx[n];
        Fy = P.Fy[n];
    }
    float ft[Q9], fb[Q9];
#pragma unroll
    for (int q = 0; q < Q9; ++q) ft[q] = P.f[q * N + n];
    pp_barplus(k, (float)P.rho[n], P.U[n], P.V[n], P.tau[n], Fx, Fy, ft, fb);
#pragma unroll
    for (int q = 0; q < Q9; ++q) P.fbp[q * N + n] = fb[q];
}

// CD face value of population q (CD_Interior_phi_Bh, DUGKS2DSolver.cpp:604-687)
// on face t of direction dir at (i, j): owner west / south, neighbour east / north
__device__ __forceinline__ float pp32_face_value(const PP32Planes &P, const KPP &k, int dir, int q, int i, int j)
{
    const int nx = P.nx, ny = P.ny;
    const float *b = P.fbp + (size_t)q * P.N;
    auto B = [&](int a, int c) { return b[(size_t)wrapp(c, ny) * nx + wrapp(a, nx)]; };
    if (dir == 0) {
        const int oi = i - 1;
        return pp_recon<0>(k, q, B(oi, j), B(i, j), B(i, j + 1), B(i, j - 1), B(oi, j + 1), B(oi, j - 1));
    }
    const int oj = j - 1;
    return pp_recon<1>(k, q, B(i, oj), B(i, j), B(i + 1, j), B(i - 1, j), B(i + 1, oj), B(i - 1, oj));
}

// face macros (Update_MacroVar_h, D2Q9PP.cpp:110-136) from the nine face
// values and the face's force of the previous step: rho, U, V
__device__ __forceinline__ void pp_face_macros(const KPP &k, const float *fb, float ffx, float ffy, float &rho,
                                               float &U, float &V)
{
    float s = 0.f, sx = 0.f, sy = 0.f;
#pragma unroll
    for (int q = 0; q < Q9; ++q) s += fb[q];
#pragma unroll
    for (int q = 0; q < Q9; ++q) sx += (float)ex9(q) * fb[q];
#pragma unroll
    for (int q = 0; q < Q9; ++q) sy += (float)ey9(q) * fb[q];
    rho = s;
    const float r = 1.f / s;
    U = (k.ms * sx + k.half_hdt * ffx) * r;
    V = (k.ms * sy + k.half_hdt * ffy) * r;
}

// half-step BGK of population q at a face (Update_phi_Eqh / Update_DVDF_Source_h /
// Update_phi_h, D2Q9PP.cpp:63-74,146-154, DUGKS2DSolver.cpp:402-423)
__device__ __forceinline__ float pp_face_bgk(const KPP &k, int q, float fb, float rho, float U, float V, float uul,
                                             float Fx, float Fy, float irr)
{
    float fso;
    const float feq = pp_feq(k, q, rho, U, V, uul, Fx, Fy, irr, fso);
    return k.ah * fb + k.bh * feq + k.ch * fso;
}

// cell velocity (Update_MacroVar, D2Q9PP.cpp:89-109) from the new populations,
// the density and the cell force: u = sum xi f~ / rho + hdt F / rho
__device__ __forceinline__ void pp_cell_macros(const KPP &k, const float *ft, float rho, float Fx, float Fy, float &U,
                                               float &V)
{
    float sx = 0.f, sy = 0.f;
#pragma unroll
    for (int q = 0; q < Q9; ++q) sx += (float)ex9(q) * ft[q];
#pragma unroll
    for (int q = 0; q < Q9; ++q) sy += (float)ey9(q) * ft[q];
    const float r = 1.f / rho;
    U = (k.ms * sx + k.hdt * Fx) * r;
    V = (k.ms * sy + k.hdt * Fy) * r;
}

// faces, pass 1: face macros (Update_MacroVar_h, D2Q9PP.cpp:110-136) with the
// face force of the previous step, and Psi of the face (Update_PseudoPsi,
// PseudoP.cpp:23-30) in FP64
__global__ void __launch_bounds__(256) k_pp32_face_psi(PP32Planes P, Consts C, KPP k)
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
        fb[q] = pp32_face_value(P, k, dir, q, i, j);
        fv[(size_t)q * N + fi] = fb[q];  // the flux pass reads it back and overwrites it with the flux
    }
    float rho, U, V, scg, psi;
    pp_face_macros(k, fb, P.ffx[t], P.ffy[t], rho, U, V);
    pp_psi(C, (double)rho, scg, psi);
    P.fpsi[t] = psi;
    P.frho[t] = rho;
    P.fU[t] = U;
    P.fV[t] = V;
    P.fscg[t] = scg;
}

// faces, pass 2: face force from the face-lattice gradient of Psi
// (update_MQ_x/y(Face_2D*), GradSchemeBasic.h:181-208; Update_PseudoForce(face),
// PseudoP.cpp:39-46), half-step BGK (DUGKS2DSolver.cpp:402-423) and micro-flux
// (:424-441) of the face values that pass 1 left in the flux planes
__global__ void __launch_bounds__(256) k_pp32_face_flux(PP32Planes P, KPP k)
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
    pp_grad9(k, c, d, gx, gy);
    const float sp = -P.fscg[t] * P.fpsi[t];
    gx *= sp;
    gy *= sp;
    P.ffx[t] = gx;
    P.ffy[t] = gy;
    const float rho = P.frho[t], U = P.fU[t], V = P.fV[t];
    const float uul = (U * U + V * V) * k.lambda0, irr = k.inv_rt / rho;
    float *fl = P.ffl + (dir ? (size_t)Q9 * N : 0);
    float mass = 0.f;
#pragma unroll
    for (int q = 0; q < Q9; ++q) {
        const int en = dir == 0 ? ex9(q) : ey9(q);
        if (en == 0) {  // no flux
            fl[(size_t)q * N + fi] = 0.f;
            continue;
        }
        const float ff = pp_face_bgk(k, q, fl[(size_t)q * N + fi], rho, U, V, uul, gx, gy, irr);
        const float flx = en > 0 ? ff * k.flux : -(ff * k.flux);
        fl[(size_t)q * N + fi] = flx;
        mass += flx;
    }
    P.fmass[t] = mass;
}

// Update_phi_T (DUGKS2DSolver.cpp:748-792), Update_MacroVar (D2Q9PP.cpp:89-109)
// and Psi of the cell (FP64 EOS)
__global__ void __launch_bounds__(256) k_pp32_update(PP32Planes P, Consts C, KPP k)
{
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t N = P.N;
    if (n >= N) return;
    const int nx = P.nx, ny = P.ny;
    const int i = (int)(n % nx), j = (int)(n / nx);
    const size_t kE = (size_t)j * nx + wrapp(i + 1, nx), kW = n;
    const size_t kN = (size_t)wrapp(j + 1, ny) * nx + i, kS = n;
    const float aTP = 4.f / 3.f, bTP = 1.f / 3.f, dtv = k.dtv;
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
    float U, V, scg, psi;
    pp_cell_macros(k, ft, (float)rho_d, P.Fx[n], P.Fy[n], U, V);
    pp_psi(C, rho_d, scg, psi);
    P.rho[n] = rho_d;
    P.U[n] = U;
    P.V[n] = V;
    P.psi[n] = psi;
    P.scg[n] = scg;
}

// the cell force of the current state (the last phase of the reference step),
// for downloads: the steps themselves compute it at the start of the next step
__global__ void __launch_bounds__(256) k_pp32_force(PP32Planes P, KPP k)
{
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (size_t)P.N) return;
    float Fx, Fy;
    pp32_cell_force(P, k, n, Fx, Fy);
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

// ---------------- one launch per step ----------------
//
// One 16x15 tile per 256-thread workgroup does the four phases above for its
// cells, reading the current set with halos and writing the other set:
//  S0: Psi and scG on the tile plus three cells (the cell force of a bar-plus cell);
//  S1: cell force and bar-plus on the tile plus two cells (the cells a face-Psi face reads);
//  S2: face macros and Psi on the x- and y-faces within one face of the tile's faces
//      (the tile's own faces also keep their nine face values);
//  S3: face force, BGK and flux of the tile's own faces (17x15 x-faces, 16x16 y-faces);
//  S4: update of the tile's cells.
// Every value is computed with the same device functions as the per-phase
// kernels, from the same inputs.
// Tile shapes: 16x15 cells on 256 threads, 32x16 on 512 (per cell about 10%
// less halo work, and twice the LDS: 81.3 KB, two workgroups per CU).
// H: halo of the face values' cells beyond the faces' own (1 CD; 2 UW, whose
// 6-point gradients reach one cell further)
template <int FTX_, int FTY_, int H_ = 1>
struct PPTile {
    static constexpr int FTX = FTX_, FTY = FTY_, H = H_, FNT = FTX * FTY > 256 ? 512 : 256;
    static constexpr int PO = 2 + H, RO = 1 + H;                             // region offsets
    static constexpr int QX = FTX + 2 * PO, QY = FTY + 2 * PO, QN = QX * QY; // Psi region: -PO .. +PO-1
    static constexpr int RX = FTX + 2 * RO, RY = FTY + 2 * RO, RN = RX * RY; // bar-plus region: -RO .. +RO-1
    static constexpr int XPX = FTX + 3, XPY = FTY + 2, XPN = XPX * XPY;      // x-face Psi region: fx -1..TX+1, fy -1..TY
    static constexpr int YPX = FTX + 2, YPY = FTY + 3, YPN = YPX * YPY;      // y-face Psi region: fx -1..TX, fy -1..TY+1
    static constexpr int XFX = FTX + 1, XFN = XFX * FTY;                     // flux x-faces: fx 0..TX, fy 0..TY-1
    static constexpr int YFN = FTX * (FTY + 1);                              // flux y-faces: fx 0..TX-1, fy 0..TY
    static constexpr int FN = XFN + YFN;
    static_assert(FTX * FTY <= FNT, "one tile cell per thread in the update");
};
// the six populations with a non-zero normal velocity on x- and y-faces (the
// other three have no flux), in increasing q; slot of q, or -1
constexpr int MX[6] = {1, 2, 4, 5, 6, 8}, MY[6] = {2, 3, 4, 6, 7, 8};
__host__ __device__ constexpr int slot_x(int q) { return q == 1 ? 0 : q == 2 ? 1 : q == 4 ? 2 : q == 5 ? 3 : q == 6 ? 4 : q == 8 ? 5 : -1; }
__host__ __device__ constexpr int slot_y(int q) { return q == 2 ? 0 : q == 3 ? 1 : q == 4 ? 2 : q == 6 ? 3 : q == 7 ? 4 : q == 8 ? 5 : -1; }

__device__ __forceinline__ int wrapn(int i, int n)
{
    const int r = i % n;
    return r < 0 ? r + n : r;
}

__device__ __forceinline__ int fused_tile_of_block(int b, int ntiles)
{
    const int per = (ntiles + 7) / 8;
    const int t = (b % 8) * per + b / 8;  // XCD k (block b mod 8) takes a contiguous band of tiles
    return t < ntiles ? t : -1;
}

// CD face value of population q from the bar-plus region (index of cell (cx, cy))
template <class G>
__device__ __forceinline__ float fused_face_value(const KPP &k, const float *bp, int dir, int q, int fx, int fy)
{
    constexpr int RX = G::RX, RO = G::RO;
    auto B = [&](int cx, int cy) { return bp[(cy + RO) * RX + (cx + RO)]; };
    if (dir == 0) {
        const int ox = fx - 1;
        return pp_recon<0>(k, q, B(ox, fy), B(fx, fy), B(fx, fy + 1), B(fx, fy - 1), B(ox, fy + 1), B(ox, fy - 1));
    }
    const int oy = fy - 1;
    return pp_recon<1>(k, q, B(fx, oy), B(fx, fy), B(fx + 1, fy), B(fx - 1, fy), B(fx + 1, oy), B(fx - 1, oy));
}

// UW face value of population q (UW_Interior_phi_Bh + Update_phi_fBh,
// DUGKS2DSolver.cpp:560-603,688-706; gradients Grad_VS_6points,
// GradSchemeBasic.h:95-126) from the bar-plus region: the upwind cell's value
// plus its 6-point gradient along x_f - h xi - x_c; for e_n = 0 the mean of
// both cells' (owner west / south, neighbour east / north)
template <class G>
__device__ __forceinline__ float fused_face_value_uw(const KPP &k, const float *bp, int dir, int q, int fx, int fy)
{
    constexpr int RX = G::RX, RO = G::RO;
    auto B = [&](int cx, int cy) { return bp[(cy + RO) * RX + (cx + RO)]; };
    const int en = dir == 0 ? ex9(q) : ey9(q), et = dir == 0 ? ey9(q) : ex9(q);
    const float ct = et == 0 ? 0.f : (et > 0 ? -k.uw_t : k.uw_t);
    auto side = [&](int cx, int cy, float cn) {
        const float e = B(cx + 1, cy), w = B(cx - 1, cy), n = B(cx, cy + 1), s = B(cx, cy - 1);
        const float ne = B(cx + 1, cy + 1), nw = B(cx - 1, cy + 1), sw = B(cx - 1, cy - 1), se = B(cx + 1, cy - 1);
        const float gx = 4.f * (e - w) + ((ne - nw) + (se - sw));
        const float gy = 4.f * (n - s) + ((ne - se) + (nw - sw));
        float v = fmaf(cn, dir == 0 ? gx : gy, B(cx, cy));
        if (et != 0) v = fmaf(ct, dir == 0 ? gy : gx, v);
        return v;
    };
    const int ox = dir == 0 ? fx - 1 : fx, oy = dir == 0 ? fy : fy - 1;
    if (en > 0) return side(ox, oy, k.uw_n);
    if (en < 0) return side(fx, fy, -k.uw_n);
    return 0.5f * (side(ox, oy, k.uw_0) + side(fx, fy, -k.uw_0));
}

template <bool FORCE, int TX, int TY, int FLUX = MPD_FLUX_CD>
__global__ void __launch_bounds__(TX * TY > 256 ? 512 : 256) k_pp32_tile(PP32Planes in, PP32Set out, Consts C, KPP k,
                                                                        int ntx, int ntiles)
{
    using G = PPTile<TX, TY, FLUX == MPD_FLUX_UW ? 2 : 1>;
    constexpr int FTX = G::FTX, FTY = G::FTY, FNT = G::FNT, QX = G::QX, QN = G::QN, RX = G::RX, RN = G::RN;
    constexpr int PO = G::PO, RO = G::RO;
    constexpr int XPX = G::XPX, XPN = G::XPN, YPX = G::YPX, YPN = G::YPN, XFX = G::XFX, XFN = G::XFN, FN = G::FN;
    static_assert(FNT == (TX * TY > 256 ? 512 : 256), "launch bounds");
    __shared__ float s_bp[Q9 * RN];
    __shared__ float s_F[2][FTX * FTY];
    __shared__ float s_fm[5][XPN + YPN];   // face Psi, scG, rho, U, V
    __shared__ float s_fl[6 * FN];         // face values of the tile's faces, then their fluxes
    // cell Psi and scG (S0-S1) live in the face-macro array until S2 writes it
    float *const s_psi = &s_fm[0][0], *const s_scg = s_psi + QN;
    static_assert(2 * QN <= 5 * (XPN + YPN), "Psi staging must fit in the face-macro array");
    const int tile = fused_tile_of_block(blockIdx.x, ntiles);
    if (tile < 0) return;
    const int nx = in.nx, ny = in.ny;
    const size_t N = in.N;
    const int i0 = (tile % ntx) * FTX, j0 = (tile / ntx) * FTY;
    const int tid = threadIdx.x;

    // S0: Psi, scG
    for (int q = tid; q < QN; q += FNT) {
        const int cx = q % QX - PO, cy = q / QX - PO;
        const size_t n = (size_t)wrapn(j0 + cy, ny) * nx + wrapn(i0 + cx, nx);
        s_psi[q] = in.psi[n];
        s_scg[q] = in.scg[n];
    }
    __syncthreads();

    // S1: cell force and bar-plus
    for (int r = tid; r < RN; r += FNT) {
        const int cx = r % RX - RO, cy = r / RX - RO;
        const size_t n = (size_t)wrapn(j0 + cy, ny) * nx + wrapn(i0 + cx, nx);
        float Fx, Fy;
        if (FORCE) {
            auto S = [&](int a, int b) { return s_psi[(b + PO) * QX + (a + PO)]; };
            const float c[4] = {S(cx + 1, cy), S(cx, cy + 1), S(cx - 1, cy), S(cx, cy - 1)};
            const float d[4] = {S(cx + 1, cy + 1), S(cx - 1, cy + 1), S(cx - 1, cy - 1), S(cx + 1, cy - 1)};
            float gx, gy;
            pp_grad9(k, c, d, gx, gy);
            const int qi = (cy + PO) * QX + (cx + PO);
            const float sp = -s_scg[qi] * s_psi[qi];
            Fx = gx * sp;
            Fy = gy * sp;
        } else {
            Fx = in.Fx[n];
            Fy = in.Fy[n];
        }
        if (cx >= 0 && cx < FTX && cy >= 0 && cy < FTY) {
            s_F[0][cy * FTX + cx] = Fx;
            s_F[1][cy * FTX + cx] = Fy;
        }
        float ft[Q9], fb[Q9];
#pragma unroll
        for (int q = 0; q < Q9; ++q) ft[q] = in.f[q * N + n];
        pp_barplus(k, (float)in.rho[n], in.U[n], in.V[n], in.tau[n], Fx, Fy, ft, fb);
#pragma unroll
        for (int q = 0; q < Q9; ++q) s_bp[q * RN + r] = fb[q];
    }
    __syncthreads();

    // S2: face macros with the face's previous force, and face Psi (FP64 EOS)
    for (int kk = tid; kk < XPN + YPN; kk += FNT) {
        const int dir = kk < XPN ? 0 : 1;
        int fx, fy;
        if (dir == 0) { fx = kk % XPX - 1; fy = kk / XPX - 1; }
        else { fx = (kk - XPN) % YPX - 1; fy = (kk - XPN) / YPX - 1; }
        const size_t gf = (dir ? N : 0) + (size_t)wrapn(j0 + fy, ny) * nx + wrapn(i0 + fx, nx);
        float fb[Q9];
#pragma unroll
        for (int q = 0; q < Q9; ++q) {
            if constexpr (FLUX == MPD_FLUX_UW) fb[q] = fused_face_value_uw<G>(k, s_bp + q * RN, dir, q, fx, fy);
            else fb[q] = fused_face_value<G>(k, s_bp + q * RN, dir, q, fx, fy);
        }
        // a face of the tile keeps its values for the flux stage
        int kf = -1;
        if (dir == 0 && fx >= 0 && fx <= FTX && fy >= 0 && fy < FTY) kf = fy * XFX + fx;
        if (dir == 1 && fx >= 0 && fx < FTX && fy >= 0 && fy <= FTY) kf = XFN + fy * FTX + fx;
        if (kf >= 0) {
#pragma unroll
            for (int z = 0; z < 6; ++z) s_fl[z * FN + kf] = fb[dir ? MY[z] : MX[z]];
        }
        float rho, U, V, scg, psi;
        pp_face_macros(k, fb, in.ffx[gf], in.ffy[gf], rho, U, V);
        pp_psi(C, (double)rho, scg, psi);
        s_fm[0][kk] = psi;
        s_fm[1][kk] = scg;
        s_fm[2][kk] = rho;
        s_fm[3][kk] = U;
        s_fm[4][kk] = V;
    }
    __syncthreads();

    // S3: face force, half-step BGK and micro-flux of the tile's faces
    for (int kk = tid; kk < FN; kk += FNT) {
        const int dir = kk < XFN ? 0 : 1;
        int fx, fy, pk;
        if (dir == 0) { fx = kk % XFX; fy = kk / XFX; pk = (fy + 1) * XPX + (fx + 1); }
        else { fx = (kk - XFN) % FTX; fy = (kk - XFN) / FTX; pk = XPN + (fy + 1) * YPX + (fx + 1); }
        const int pw = dir ? YPX : XPX;
        auto S = [&](int a, int b) { return s_fm[0][pk + b * pw + a]; };
        const float c[4] = {S(1, 0), S(0, 1), S(-1, 0), S(0, -1)};
        const float d[4] = {S(1, 1), S(-1, 1), S(-1, -1), S(1, -1)};
        float gx, gy;
        pp_grad9(k, c, d, gx, gy);
        const float sp = -s_fm[1][pk] * s_fm[0][pk];
        gx *= sp;
        gy *= sp;
        const bool own = (dir == 0 ? fx < FTX : fy < FTY) && i0 + fx < nx && j0 + fy < ny;
        if (own) {
            const size_t gf = (dir ? N : 0) + (size_t)(j0 + fy) * nx + (i0 + fx);
            out.ffx[gf] = gx;
            out.ffy[gf] = gy;
        }
        const float rho = s_fm[2][pk], U = s_fm[3][pk], V = s_fm[4][pk];
        const float uul = (U * U + V * V) * k.lambda0, irr = k.inv_rt / rho;
        // the populations with a zero normal velocity carry no flux
        if (dir == 0) {
#pragma unroll
            for (int z = 0; z < 6; ++z) {
                const float ff = pp_face_bgk(k, MX[z], s_fl[z * FN + kk], rho, U, V, uul, gx, gy, irr);
                s_fl[z * FN + kk] = ex9(MX[z]) > 0 ? ff * k.flux : -(ff * k.flux);
            }
        } else {
#pragma unroll
            for (int z = 0; z < 6; ++z) {
                const float ff = pp_face_bgk(k, MY[z], s_fl[z * FN + kk], rho, U, V, uul, gx, gy, irr);
                s_fl[z * FN + kk] = ey9(MY[z]) > 0 ? ff * k.flux : -(ff * k.flux);
            }
        }
    }
    __syncthreads();

    // S4: Update_phi_T, the conserved FP64 density, Update_MacroVar, Psi
    const int cx = tid % FTX, cy = tid / FTX;
    if (tid >= FTX * FTY || i0 + cx >= nx || j0 + cy >= ny) return;
    const size_t n = (size_t)(j0 + cy) * nx + (i0 + cx);
    const int kW = cy * XFX + cx, kE = kW + 1;
    const int kS = XFN + cy * FTX + cx, kN = kS + FTX;
    const int rc = (cy + RO) * RX + (cx + RO);
    const float aTP = 4.f / 3.f, bTP = 1.f / 3.f, dtv = k.dtv;
    float ft[Q9];
    double rest = 0.0;
#pragma unroll
    for (int q = 0; q < Q9; ++q) {
        const int zx = slot_x(q), zy = slot_y(q);
        float s = 0.f;
        if (zx >= 0) s += -s_fl[zx * FN + kE];
        if (zy >= 0) s += -s_fl[zy * FN + kN];
        if (zx >= 0) s += s_fl[zx * FN + kW];
        if (zy >= 0) s += s_fl[zy * FN + kS];
        ft[q] = aTP * s_bp[q * RN + rc] - bTP * in.f[q * N + n] + dtv * s;
        if (q > 0) rest += (double)ft[q];
    }
    // each face's mass flux: the sum of its population fluxes in increasing q
    auto mass = [&](int kf) {
        float m = 0.f;
#pragma unroll
        for (int z = 0; z < 6; ++z) m += s_fl[z * FN + kf];
        return m;
    };
    const double dm = ((double)mass(kW) - (double)mass(kE)) + ((double)mass(kS) - (double)mass(kN));
    const double rho_d = in.rho[n] + C.dtv * dm;
    ft[0] = (float)(rho_d - rest);
#pragma unroll
    for (int q = 0; q < Q9; ++q) out.f[q * N + n] = ft[q];
    float U, V, scg, psi;
    pp_cell_macros(k, ft, (float)rho_d, s_F[0][cy * FTX + cx], s_F[1][cy * FTX + cx], U, V);
    pp_psi(C, rho_d, scg, psi);
    out.rho[n] = rho_d;
    out.U[n] = U;
    out.V[n] = V;
    out.psi[n] = psi;
    out.scg[n] = scg;
}

}  // namespace

int pp32_alloc(PP32State &St, void **mem, int64_t *bytes, const Consts &C, int64_t N)
{
    // two sets of: FP64 density [N]; FP32 f 9, U V Fx Fy Psi scG 6 [N], face
    // force 2 [2N]; then tau [N] and the per-phase scratch: bar-plus 9 [N], face
    // Psi rho U V scG mass 6 [2N], fluxes 9 [2N]
    const size_t n = (size_t)N;
    const size_t set_f = 15 * n + 4 * n, scratch_f = n + 9 * n + 12 * n + 18 * n;
    const size_t bytes_total = 2 * n * sizeof(double) + (2 * set_f + scratch_f) * sizeof(float);
    void *base = nullptr;
    if (hipMalloc(&base, bytes_total) != hipSuccess) return -1;
    *mem = base;
    *bytes = (int64_t)bytes_total;
    PP32Planes &P = St.P;
    P.nx = C.nx;
    P.ny = C.ny;
    P.N = N;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
        St.cus = cus;
    double *d = (double *)base;
    float *p = (float *)(d + 2 * n);
    PP32Set sets[2];
    for (int w = 0; w < 2; ++w) {
        PP32Set &S = sets[w];
        S.rho = d + w * n;
        S.f = p; p += 9 * n;
        S.U = p; p += n;
        S.V = p; p += n;
        S.Fx = p; p += n;
        S.Fy = p; p += n;
        S.psi = p; p += n;
        S.scg = p; p += n;
        S.ffx = p; p += 2 * n;
        S.ffy = p; p += 2 * n;
    }
    P.rho = sets[0].rho;
    P.f = sets[0].f;
    P.U = sets[0].U;
    P.V = sets[0].V;
    P.Fx = sets[0].Fx;
    P.Fy = sets[0].Fy;
    P.psi = sets[0].psi;
    P.scg = sets[0].scg;
    P.ffx = sets[0].ffx;
    P.ffy = sets[0].ffy;
    St.other = sets[1];
    P.tau = p; p += n;
    P.fbp = p; p += 9 * n;
    P.fpsi = p; p += 2 * n;
    P.frho = p; p += 2 * n;
    P.fU = p; p += 2 * n;
    P.fV = p; p += 2 * n;
    P.fscg = p; p += 2 * n;
    P.fmass = p; p += 2 * n;
    P.ffl = p;
    return 0;
}

void pp32_from64(const PP32State &St, const DevState &S, hipStream_t st)
{
    hipLaunchKernelGGL(k_pp32_from64, dim3(blocks(St.P.N)), dim3(256), 0, st, St.P, S);
}

void pp32_to64(const PP32State &St, const DevState &S, const Consts &C, bool force_stale, hipStream_t st)
{
    if (force_stale) hipLaunchKernelGGL(k_pp32_force, dim3(blocks(St.P.N)), dim3(256), 0, st, St.P, make_kpp(C));
    hipLaunchKernelGGL(k_pp32_to64, dim3(blocks(St.P.N)), dim3(256), 0, st, St.P, S, C);
}

int pp32_tile_auto(int nx, int ny, int cus)
{
    // 32x16 tiles where every CU gets at least one, else 16x15 (at 512x512:
    // 18.8 against 26.3 us per step; 16x15 makes 1,120 tiles, a partly filled
    // second round of the 1,024 resident workgroups)
    return (int64_t)((nx + 31) / 32) * ((ny + 15) / 16) >= (cus > 0 ? cus : 256) ? MPD_TILE_32X16 : MPD_TILE_16X15;
}

template <int TX, int TY, int FLUX = MPD_FLUX_CD>
void pp32_tile_launch(PP32Planes &P, PP32State &St, const Consts &C, const KPP &k, bool force_stale, hipStream_t st)
{
    const int ntx = (P.nx + TX - 1) / TX, nty = (P.ny + TY - 1) / TY, ntiles = ntx * nty;
    const unsigned nblocks = (unsigned)(((ntiles + 7) / 8) * 8);
    dim3 b(PPTile<TX, TY>::FNT);
    if (force_stale)
        hipLaunchKernelGGL((k_pp32_tile<true, TX, TY, FLUX>), dim3(nblocks), b, 0, st, P, St.other, C, k, ntx, ntiles);
    else
        hipLaunchKernelGGL((k_pp32_tile<false, TX, TY, FLUX>), dim3(nblocks), b, 0, st, P, St.other, C, k, ntx, ntiles);
}

void pp32_step(PP32State &St, const Consts &C, bool force_stale, hipStream_t st)
{
    PP32Planes &P = St.P;
    const size_t N = P.N;
    const KPP k = make_kpp(C);
    if (St.phases) {
        if (force_stale) hipLaunchKernelGGL(k_pp32_phase1<true>, dim3(blocks(N)), dim3(256), 0, st, P, k);
        else hipLaunchKernelGGL(k_pp32_phase1<false>, dim3(blocks(N)), dim3(256), 0, st, P, k);
        hipLaunchKernelGGL(k_pp32_face_psi, dim3(blocks(2 * N)), dim3(256), 0, st, P, C, k);
        hipLaunchKernelGGL(k_pp32_face_flux, dim3(blocks(2 * N)), dim3(256), 0, st, P, k);
        hipLaunchKernelGGL(k_pp32_update, dim3(blocks(N)), dim3(256), 0, st, P, C, k);
        return;
    }
    const int tile = St.tile ? St.tile : pp32_tile_auto(P.nx, P.ny, St.cus);
    if (C.flux == MPD_FLUX_UW) pp32_tile_launch<16, 15, MPD_FLUX_UW>(P, St, C, k, force_stale, st);  // 16x15 only
    else if (tile == MPD_TILE_32X16) pp32_tile_launch<32, 16>(P, St, C, k, force_stale, st);
    else pp32_tile_launch<16, 15>(P, St, C, k, force_stale, st);
    // the written set becomes the current one
    PP32Set cur{P.rho, P.f, P.U, P.V, P.Fx, P.Fy, P.psi, P.scg, P.ffx, P.ffy};
    const PP32Set &o = St.other;
    P.rho = o.rho;
    P.f = o.f;
    P.U = o.U;
    P.V = o.V;
    P.Fx = o.Fx;
    P.Fy = o.Fy;
    P.psi = o.psi;
    P.scg = o.scg;
    P.ffx = o.ffx;
    P.ffy = o.ffy;
    St.other = cur;
}

}  // namespace mpd
