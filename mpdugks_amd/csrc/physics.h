// Device physics of the 2D DUGKS step, shared by the phase kernels and the
// fused step kernels.  Every function restates one reference routine and keeps
// its floating-point operation order (compile with -ffp-contract=off to get
// the reference's FP64 results bit for bit).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mpdugks.h"

namespace mpd {

constexpr int Q9 = 9;

// D2Q9 directions (Ark/Lib/D2Q9.h:24-28), weights (Ark/Lib/D2Q9HCZ.cpp:15-19)
__host__ __device__ constexpr int ex9(int q) { return q == 0 ? 0 : (q == 1 || q == 2 || q == 8) ? 1 : (q == 3 || q == 7) ? 0 : -1; }
__host__ __device__ constexpr int ey9(int q) { return (q == 0 || q == 1 || q == 5) ? 0 : (q >= 2 && q <= 4) ? 1 : -1; }
__host__ __device__ constexpr int bb9(int q) { return q == 0 ? 0 : (q <= 4 ? q + 4 : q - 4); }

// Constants of one case, derived on the host from mpd_config with the same
// expressions as Ark/Constant/ZeroNamespace.h / ZeroReference.h.
struct Consts {
    int32_t nx, ny, model, flux, walls, nyf;
    int32_t nee;      // walls use non-equilibrium extrapolation (Wall_3_NEE) instead of bounce-back
    int32_t unit;     // dx a power of two and xi exactly the integer directions (MaSpan = 1)
    int32_t face_ref; // pseudopotential faces in the reference's one-thread order (mpd_set_face_order)
    double dx, dt, hdt, rt, lambda0, aphi, bphi, beta, kappa, kp, taum, tau0;
    double phi_l, phi_v, nu_l, nu_v, gx, gy, rho_ref, pp_k, t_cs, pp_tau;
    double dtv;       // dt / volume                    (Cell_2D.cpp:182)
    double d2x6;      // (2*dx)*6, gradient denominator (GradSchemeBasic.h:165)
    double lapden;    // 6*MinL*MinL                    (GradScheme.cpp:362)
    double rtdt;      // RT*dt                          (SourceM.cpp:51)
    double wdphi;     // wI*(PhiL-PhiV)                 (SourceM.cpp:17)
    double twodx;     // 2*dx, CD tangential span       (MeshCartesian.cpp:145,203)
    double rdx, rtwodx;  // 1/dx, 1/(2*dx): exact reciprocals when unit
    double ah_h, bh_h, ch_h, abp_h, bbp_h, cbp_h;      // constant tau_h factors
    double ah_pp, bh_pp, ch_pp;                         // PP face factors
    double xi[2][Q9];
    double w[Q9];
};

// Carnahan-Starling EOS, a = 1, b = 4 (Ark/Lib/nonidealEoS.h:59-73).  Always FP64.
__host__ __device__ inline double carnahan_starling(double rho, double t_cs)
{
    return rho * 1.0 * t_cs * (1.0 + rho + rho * rho - rho * rho * rho) /
               ((1.0 - rho) * (1.0 - rho) * (1.0 - rho)) -
           1.0 * rho * rho;
}

// MacroQuantity::calcPsi (Ark/Header/Macroscopic.h:64-68)
__host__ __device__ inline void calc_psi(double p, double rho, double rt, double &scg, double &psi)
{
    double s = (double)((p > rho * rt) - (p < rho * rt));
    scg = s;
    psi = sqrt(2 * (p - rho * rt) / s);
}

// Cell macros consumed by phase 1.
template <typename T>
struct CellMacro {
    T phi, rho, U, V, p, tau, px, py, rx, ry, Fx, Fy;
};

// Lattice velocity component d of population q.  UNIT (Consts::unit: MaSpan
// exactly 1) gives the integer direction as a constant: the multiplications by
// +-1 then fold away, which is exact, and those by 0 stay (IEEE does not let
// the compiler drop them), so every result is the same bit for bit.
template <bool UNIT, typename T>
__device__ inline T xi_of(const Consts &C, int d, int q)
{
    return UNIT ? (T)(d == 0 ? ex9(q) : ey9(q)) : (T)C.xi[d][q];
}

// Equilibria and sources of population q of one cell for the Allen-Cahn
// models (the expressions of cell_barplus below).
template <typename T, int MODEL, bool UNIT = false>
__device__ inline void cell_eq_so(const Consts &C, const CellMacro<T> &m, int q, T &heq, T &feq, T &hso, T &fso)
{
    const T rt = (T)C.rt;
    const T uu = m.U * m.U + m.V * m.V;
    const T ex = xi_of<UNIT, T>(C, 0, q), ey = xi_of<UNIT, T>(C, 1, q), w = (T)C.w[q];
    T u1 = (ex * m.U + ey * m.V) / rt;
    T G = u1 + (T)0.5 * u1 * u1 - uu * (T)C.lambda0;
    heq = w * ((T)1.0 + u1) * m.phi;
    hso = (m.px * ex + m.py * ey) * w;
    if (MODEL == MPD_MODEL_AC_HCZ) {
        feq = w * (m.p + m.rho * G * rt);
        fso = w * (((T)1 + G) * ((ex - m.U) * m.Fx + (ey - m.V) * m.Fy) +
                   G * rt * ((ex - m.U) * m.rx + (ey - m.V) * m.ry));
    } else {
        feq = w * (m.p / rt + m.rho * G) - (T)(q == 0) * m.p / rt;
        fso = w * (((T)1 + G) * ((ex - m.U) * m.Fx + (ey - m.V) * m.Fy) / rt +
                   G * ((ex - m.U) * m.rx + (ey - m.V) * m.ry));
    }
}

// Update_phi_Eq + Update_DVDF_Source + Update_phi_BP for one cell
// (HCZ Ark/Lib/D2Q9HCZ.cpp:54-92,244-257; AC-p Ark/Lib/D2Q9AC.cpp:69-105,331-344;
//  PP Ark/Lib/D2Q9PP.cpp:51-62,137-145; Ark/Cpp/DUGKS2DSolver.cpp:379-401;
//  factors Ark/Cpp/Mesh/Cell_2D.cpp:120-125).
template <typename T, int MODEL, bool UNIT = false>
__device__ inline void cell_barplus(const Consts &C, const CellMacro<T> &m, const T *ht, const T *ft,
                                    T *hbp, T *fbp)
{
    const T rt = (T)C.rt;
    T abp = ((T)2.0 * m.tau - (T)C.hdt) / ((T)2.0 * m.tau + (T)C.dt);
    T bbp = (T)1.0 - abp;
    T cbp = m.tau * bbp;
    T uu = m.U * m.U + m.V * m.V;
#pragma unroll
    for (int q = 0; q < Q9; ++q) {
        const T ex = xi_of<UNIT, T>(C, 0, q), ey = xi_of<UNIT, T>(C, 1, q), w = (T)C.w[q];
        T feq, fso;
        T u1 = (ex * m.U + ey * m.V) / rt;
        if (MODEL == MPD_MODEL_PP_D2Q9) {
            feq = w * m.rho * ((T)1.0 + u1 + (T)0.5 * u1 * u1 - uu * (T)C.lambda0);
            fso = (m.Fx * (ex - m.U) + m.Fy * (ey - m.V)) * feq / (m.rho * rt);
        } else {
            T G = u1 + (T)0.5 * u1 * u1 - uu * (T)C.lambda0;
            T heq = w * ((T)1.0 + u1) * m.phi;
            T hso = (m.px * ex + m.py * ey) * w;
            if (MODEL == MPD_MODEL_AC_HCZ) {
                feq = w * (m.p + m.rho * G * rt);
                fso = w * (((T)1 + G) * ((ex - m.U) * m.Fx + (ey - m.V) * m.Fy) +
                           G * rt * ((ex - m.U) * m.rx + (ey - m.V) * m.ry));
            } else {
                feq = w * (m.p / rt + m.rho * G) - (T)(q == 0) * m.p / rt;
                fso = w * (((T)1 + G) * ((ex - m.U) * m.Fx + (ey - m.V) * m.Fy) / rt +
                           G * ((ex - m.U) * m.rx + (ey - m.V) * m.ry));
            }
            hbp[q] = (T)C.abp_h * ht[q] + (T)C.bbp_h * heq + (T)C.cbp_h * hso;
        }
        T b = abp * ft[q] + bbp * feq;
        b += cbp * fso;
        fbp[q] = b;
    }
}

// CD face value (CD_Interior_phi_Bh, Ark/Cpp/DUGKS2DSolver.cpp:604-687) on the
// structured face with owner o (low side) and neighbour n (high side):
//   normal derivative (n - o)/dx, tangential (0.5*(neigh pair) + 0.5*(owner pair))/(2dx).
template <typename T, bool UNIT = false>
__device__ inline T cd_face_value(const Consts &C, int dir, int q, T bo, T bn, T tn_hi, T tn_lo,
                                  T to_hi, T to_lo)
{
    // UNIT (dx a power of two): x * 2^-k and x / 2^k are the same real
    // number, so the correctly rounded product equals the quotient bit for bit
    const T tsum = (T)0.5 * (tn_hi - tn_lo) + (T)0.5 * (to_hi - to_lo);
    const T dn = UNIT ? (bn - bo) * (T)C.rdx : (bn - bo) / (T)C.dx;
    const T dt = UNIT ? tsum * (T)C.rtwodx : tsum / (T)C.twodx;
    const T ex = xi_of<UNIT, T>(C, 0, q), ey = xi_of<UNIT, T>(C, 1, q);
    if (dir == 0) return (T)0.5 * (bo + bn) - (T)C.hdt * (dn * ex + dt * ey);
    return (T)0.5 * (bo + bn) - (T)C.hdt * (dt * ex + dn * ey);
}

template <typename T>
struct FaceMacro {
    T phi, rho, U, V, p, tau, px, py, rx, ry, Fx, Fy;
};

// Update_MacroVar_h for the Allen-Cahn models (HCZ Ark/Lib/D2Q9HCZ.cpp:183-242,
// AC-p Ark/Lib/D2Q9AC.cpp:232-329); the six averaged sources come in already
// averaged.  Moment sums in sequential order (Ark/Lib/GaussHermite.h:1-14).
template <typename T, int MODEL, bool UNIT = false>
__device__ inline void face_macros_ac(const Consts &C, const T *hb, const T *fb, FaceMacro<T> &fm)
{
    const T rt = (T)C.rt, hdt = (T)C.hdt;
    T sh = 0;
#pragma unroll
    for (int q = 0; q < Q9; ++q) sh += hb[q];
    fm.phi = sh;
    fm.rho = (T)C.aphi * fm.phi + (T)C.bphi;
    T s = 0;
#pragma unroll
    for (int q = 0; q < Q9; ++q) s += fb[q];
    T mu;
    if (MODEL == MPD_MODEL_AC_HCZ) {
        T sx = 0, sy = 0;
#pragma unroll
        for (int q = 0; q < Q9; ++q) sx += fb[q] * xi_of<UNIT, T>(C, 0, q);
#pragma unroll
        for (int q = 0; q < Q9; ++q) sy += fb[q] * xi_of<UNIT, T>(C, 1, q);
        fm.U = sx / (fm.rho * rt);
        fm.V = sy / (fm.rho * rt);
        fm.U += (T)0.5 * hdt * fm.Fx / (fm.rho * rt);
        fm.V += (T)0.5 * hdt * fm.Fy / (fm.rho * rt);
        fm.p = s + (T)0.5 * hdt * rt * (fm.rx * fm.U + fm.ry * fm.V);
        mu = (T)C.nu_v + (fm.phi - (T)C.phi_v) * ((T)C.nu_l - (T)C.nu_v);
        mu *= fm.rho;
    } else {
        fm.U = 0;
        fm.V = 0;
        fm.p = s - fb[0];
        fm.p += ((T)0.5 * hdt * (fm.rx * fm.U + fm.ry * fm.V) -
                 (T)C.w[0] * fm.rho * (fm.U * fm.U + fm.V * fm.V) * (T)C.lambda0);
        fm.p *= (T)C.kp;
        mu = fm.rho * rt * (T)C.tau0;
    }
    fm.tau = mu / (fm.rho * rt);
}

// Update_phi_Eqh + Update_DVDF_Source_h + Update_phi_h for population q
// (HCZ Ark/Lib/D2Q9HCZ.cpp:93-131,258-271; AC-p Ark/Lib/D2Q9AC.cpp:106-144,345-358;
//  PP Ark/Lib/D2Q9PP.cpp:63-74,146-154; Ark/Cpp/DUGKS2DSolver.cpp:402-423).
template <typename T, int MODEL, bool UNIT = false>
__device__ inline void face_bgk(const Consts &C, const FaceMacro<T> &fm, int q, T hb, T fb, T ah, T bh,
                                T ch, T &hout, T &fout)
{
    const T rt = (T)C.rt;
    const T ex = xi_of<UNIT, T>(C, 0, q), ey = xi_of<UNIT, T>(C, 1, q), w = (T)C.w[q];
    T uu = fm.U * fm.U + fm.V * fm.V;
    T u1 = (ex * fm.U + ey * fm.V) / rt;
    T feq, fso;
    if (MODEL == MPD_MODEL_PP_D2Q9) {
        feq = w * fm.rho * ((T)1.0 + u1 + (T)0.5 * u1 * u1 - uu * (T)C.lambda0);
        fso = (fm.Fx * (ex - fm.U) + fm.Fy * (ey - fm.V)) * feq / (fm.rho * rt);
        hout = 0;
    } else {
        T G = u1 + (T)0.5 * u1 * u1 - uu * (T)C.lambda0;
        T heq = w * fm.phi * ((T)1.0 + u1);
        T hso = (fm.px * ex + fm.py * ey) * w;
        hout = (T)C.ah_h * hb + (T)C.bh_h * heq + (T)C.ch_h * hso;
        if (MODEL == MPD_MODEL_AC_HCZ) {
            feq = w * (fm.p + fm.rho * rt * G);
            fso = w * (((T)1 + G) * ((ex - fm.U) * fm.Fx + (ey - fm.V) * fm.Fy) +
                       G * rt * ((ex - fm.U) * fm.rx + (ey - fm.V) * fm.ry));
        } else {
            feq = w * (fm.p / rt + fm.rho * G) - (T)(q == 0) * fm.p / rt;
            fso = w * (((T)1 + G) * ((ex - fm.U) * fm.Fx + (ey - fm.V) * fm.Fy) / rt +
                       G * ((ex - fm.U) * fm.rx + (ey - fm.V) * fm.ry));
        }
    }
    T r = ah * fb + bh * feq;
    r += ch * fso;
    fout = r;
}

// Wall_3_NEE (Ark/Cpp/Boundary.cpp:140-184) for the Allen-Cahn models: face
// values of every population at the wall face (xf, yf) of fluid cell centre
// (xc, yc) with macros cm and tilde populations ht, ft.  The face macros are
// extrapolated from the cell -- Phi, Rho along the cell's Phi_x/y (the AC
// source vector at this point of the step) and Rho_x/y; p of the cell; U = V
// = 0 (the reference never sets the face velocity otherwise) -- and the
// cell's non-equilibrium part 2 tau/(2 tau + dt) (x~ - x_eq + dt/2 x_so) is
// added.  Bounce-back follows, in the caller.
template <typename T, int MODEL, bool UNIT = false>
__device__ inline void nee_face(const Consts &C, const CellMacro<T> &cm, const T *ht, const T *ft, T xf, T yf,
                                T xc, T yc, T *hh, T *ff)
{
    const T phf = cm.phi + cm.px * (xf - xc) + cm.py * (yf - yc);
    const T rhf = cm.rho + cm.rx * (xf - xc) + cm.ry * (yf - yc);
    const T pf = cm.p, Uf = 0, Vf = 0, uu = Uf * Uf + Vf * Vf;
    const T hneq = (T)2.0 * (T)C.taum / ((T)2.0 * (T)C.taum + (T)C.dt);
    const T fneq = (T)2.0 * cm.tau / ((T)2.0 * cm.tau + (T)C.dt);
    for (int q = 0; q < Q9; ++q) {
        T heq, feq, hso, fso;
        cell_eq_so<T, MODEL, UNIT>(C, cm, q, heq, feq, hso, fso);
        const T ex = xi_of<UNIT, T>(C, 0, q), ey = xi_of<UNIT, T>(C, 1, q), w = (T)C.w[q];
        const T u1 = (ex * Uf + ey * Vf) / (T)C.rt;
        const T G = u1 + (T)0.5 * u1 * u1 - uu * (T)C.lambda0;
        const T heqf = w * phf * ((T)1.0 + u1);
        const T feqf = MODEL == MPD_MODEL_AC_HCZ ? w * (pf + rhf * (T)C.rt * G)
                                                 : w * (pf / (T)C.rt + rhf * G) - (T)(q == 0) * pf / (T)C.rt;
        hh[q] = heqf + hneq * (ht[q] - heq + (T)C.hdt * hso);
        ff[q] = feqf + fneq * (ft[q] - feq + (T)C.hdt * fso);
    }
}

// Update_MacroVar for the Allen-Cahn models (HCZ Ark/Lib/D2Q9HCZ.cpp:133-182,
// AC-p Ark/Lib/D2Q9AC.cpp:146-231).  Fx/Fy/rx/ry are the sources of the
// previous MacroSource.  Writes phi, rho, U, V, p, mu, tau.
template <typename T, int MODEL, bool UNIT = false>
__device__ inline void cell_macrovar_ac(const Consts &C, const T *ht, const T *ft, T Fx, T Fy, T rx, T ry,
                                        T &phi, T &rho, T &U, T &V, T &p, T &mu, T &tau)
{
    const T rt = (T)C.rt, hdt = (T)C.hdt;
    T sh = 0;
#pragma unroll
    for (int q = 0; q < Q9; ++q) sh += ht[q];
    phi = sh;
    rho = (T)C.aphi * phi + (T)C.bphi;
    T s = 0;
#pragma unroll
    for (int q = 0; q < Q9; ++q) s += ft[q];
    if (MODEL == MPD_MODEL_AC_HCZ) {
        T sx = 0, sy = 0;
#pragma unroll
        for (int q = 0; q < Q9; ++q) sx += ft[q] * xi_of<UNIT, T>(C, 0, q);
#pragma unroll
        for (int q = 0; q < Q9; ++q) sy += ft[q] * xi_of<UNIT, T>(C, 1, q);
        U = sx / (rho * rt);
        V = sy / (rho * rt);
        U += hdt * Fx / (rho * rt);
        V += hdt * Fy / (rho * rt);
        p = s + hdt * (rx * U + ry * V) * rt;
        mu = (T)C.nu_v + (phi - (T)C.phi_v) * ((T)C.nu_l - (T)C.nu_v);
        mu *= rho;
    } else {
        U = 0;
        V = 0;
        p = s - ft[0];
        p += (hdt * (rx * U + ry * V) - rho * (T)C.w[0] * (U * U + V * V) * (T)C.lambda0);
        p *= (T)C.kp;
        mu = rho * rt * (T)C.tau0;
    }
    tau = mu / (rho * rt);
}

// 9-point isotropic gradient (update_MQ_x/y, Ark/Lib/GradSchemeBasic.h:153-180):
// c[0..3] = E,N,W,S; d[0..3] = NE,NW,SW,SE.
template <typename T>
__device__ inline void grad9(const Consts &C, const T *c, const T *d, T &gx, T &gy)
{
    gx = ((T)4 * (c[0] - c[2]) + (T)1 * (d[0] - d[1] + d[3] - d[2])) / (T)C.d2x6;
    gy = ((T)4 * (c[1] - c[3]) + (T)1 * (d[0] - d[3] + d[1] - d[2])) / (T)C.d2x6;
}

template <typename T>
__device__ inline void set_zero(T &d)
{ // SetZero, Ark/Header/Mesh_2D.h:26-30
    if (d < (T)1.0E-14 && d > (T)-1.0E-14) d = 0;
}

// Sources of MacroSource (Ark/Cpp/SourceM.cpp:15-62 with Grad_Phi_6points,
// Ark/Cpp/GradScheme.cpp:347-376).  c/d: phi of the 8 neighbours.
template <typename T>
struct AcSource {
    T lap, rx, ry, Fx, Fy, sx, sy, ppu, ppv;
};

template <typename T>
__device__ inline void ac_source(const Consts &C, T phi, T rho, T U, T V, T ppu_prev, T ppv_prev,
                                 const T *c, const T *d, AcSource<T> &s)
{
    T gx, gy;
    grad9(C, c, d, gx, gy);
    T lap = ((T)4 * (c[0] + c[1] + c[2] + c[3]) + (d[0] + d[1] + d[2] + d[3]) - (T)20 * phi) /
            (T)C.lapden;
    set_zero(gx);
    set_zero(gy);
    T mu_c = (T)2 * (T)C.beta * (phi - (T)C.phi_l) * (phi - (T)C.phi_v) *
                 ((T)2 * phi - (T)C.phi_l - (T)C.phi_v) -
             (T)C.kappa * lap;
    s.rx = (T)C.aphi * gx;
    s.ry = (T)C.aphi * gy;
    s.Fx = mu_c * gx;
    s.Fy = mu_c * gy;
    if (C.gx != 0.0) s.Fx -= (T)C.gx * (rho - (T)C.rho_ref);
    if (C.gy != 0.0) s.Fy -= (T)C.gy * (rho - (T)C.rho_ref);
    T dpu = phi * U - ppu_prev;
    T dpv = phi * V - ppv_prev;
    T L = sqrt(gx * gx + gy * gy);
    if (L != (T)0) {
        T tmp = (T)-4 * (phi - (T)C.phi_l) * (phi - (T)C.phi_v) / (T)C.wdphi;
        s.sx = gx * tmp / L + dpu / (T)C.rtdt;
        s.sy = gy * tmp / L + dpv / (T)C.rtdt;
    } else {
        s.sx = dpu / (T)C.rtdt;
        s.sy = dpv / (T)C.rtdt;
    }
    s.lap = lap;
    s.ppu = phi * U;
    s.ppv = phi * V;
}

}  // namespace mpd
