// FP32-bulk fused DUGKS step (Allen-Cahn models, CD faces, periodic x,
// periodic or bounce-back y).  Same stages and derived state as the FP64
// fused kernel (fused_kernels.hip), arithmetic rewritten for throughput:
//  - D2Q9 directions and weights are compile-time constants, so e_q . u
//    folds to adds and the populations with zero normal velocity skip the
//    face BGK (their flux is zero);
//  - divisions by per-cell/per-face scalars become one reciprocal each;
//  - the three convex combinations are written as  x + b (y - x) + c z  so the
//    unit coefficient is exact and FP32 rounding of a and b = 1 - a cannot make
//    the scheme gain or lose mass;
//  - the D2Q9 weights are rounded so that they sum to exactly 1 in FP32;
//  - AC-p (U == 0 by construction, Ark/Lib/D2Q9AC.cpp:197-198) drops every
//    velocity term and has constant relaxation factors.
// f~ is stored as the deviation from the rest equilibrium at p_ref, so p below
// is p - p_ref (all equilibria and sources are linear in p).
// Compiled with FP contraction enabled.
#include <hip/hip_runtime.h>

#include <vector>

#include "fused.h"
#include "physics.h"

namespace mpd {

namespace {

constexpr int TX = 16, TY = 15;
constexpr int R1X = TX + 2, R1Y = TY + 2;
constexpr int R2X = TX + 4, R2Y = TY + 4;
constexpr int R1N = R1X * R1Y, R2N = R2X * R2Y;
constexpr int NXF = (TX + 1) * TY;  // 255
constexpr int NYF = TX * (TY + 1);  // 256
constexpr int NT = 256;
constexpr int NZ = 6;               // populations with non-zero normal velocity per face direction

// populations with e_x != 0 (x-face fluxes) and e_y != 0 (y-face fluxes)
__host__ __device__ constexpr int xq(int z) { return z == 0 ? 1 : z == 1 ? 2 : z == 2 ? 4 : z == 3 ? 5 : z == 4 ? 6 : 8; }
__host__ __device__ constexpr int yq(int z) { return z == 0 ? 2 : z == 1 ? 3 : z == 2 ? 4 : z == 3 ? 6 : z == 4 ? 7 : 8; }

struct K32 {
    float ms;        // MaSpan
    float ms_rt;     // MaSpan / RT
    float rt, inv_rt, lambda0, hdt, dt, aphi, bphi, beta, kappa, phi_l, phi_v, nu_v, dnu;
    float inv_d2x6, inv_lapden, inv_rtdt, inv_wdphi, inv_dx, inv_2dx, kp, dtv;
    float gx, gy, rho_ref;
    float bbp_h, cbp_h, bh_h, ch_h;   // Allen-Cahn factors (constant tau_h)
    float tau0;                       // AC-p constant tau
    float bbp0, cbp0, bh0, ch0;       // AC-p constant factors
    float flx;                        // MaSpan * face area
    float w[9];
};

// one case of a batched launch (device table entry)
struct BatchCase {
    K32 k;
    int nx, ny;
    int64_t N;
    FusedPlanes32 pl[2];
    int cur;  // plane set holding the current state
};

__device__ __forceinline__ int wrapc(int i, int n)
{
    int r = i % n;
    return r < 0 ? r + n : r;
}

__device__ __forceinline__ int tile_of_block(int b, int ntiles)
{
    const int nxcd = 8;
    int per = (ntiles + nxcd - 1) / nxcd;
    int t = (b % nxcd) * per + b / nxcd;
    return t < ntiles ? t : -1;
}

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// grad(phi), laplacian, SetZero (GradScheme.cpp:347-376); c: E N W S, d: NE NW SW SE
__device__ __forceinline__ void stencil(const K32 &k, float phi, const float *c, const float *d, float &gx, float &gy,
                                        float &lap)
{
    gx = (4.f * (c[0] - c[2]) + (d[0] - d[1] + d[3] - d[2])) * k.inv_d2x6;
    gy = (4.f * (c[1] - c[3]) + (d[0] - d[3] + d[1] - d[2])) * k.inv_d2x6;
    lap = (4.f * (c[0] + c[1] + c[2] + c[3]) + (d[0] + d[1] + d[2] + d[3]) - 20.f * phi) * k.inv_lapden;
    if (fabsf(gx) < 1.0E-14f) gx = 0.f;
    if (fabsf(gy) < 1.0E-14f) gy = 0.f;
}

// SourceMomentum (SourceM.cpp:19-34)
__device__ __forceinline__ void momentum_source(const K32 &k, float phi, float gx, float gy, float lap, float &Fx,
                                                float &Fy, float &rx, float &ry)
{
    float mu_c = 2.f * k.beta * (phi - k.phi_l) * (phi - k.phi_v) * (2.f * phi - k.phi_l - k.phi_v) - k.kappa * lap;
    rx = k.aphi * gx;
    ry = k.aphi * gy;
    Fx = mu_c * gx;
    Fy = mu_c * gy;
    float rho = k.aphi * phi + k.bphi;
    Fx -= k.gx * (rho - k.rho_ref);
    Fy -= k.gy * (rho - k.rho_ref);
}

// One face: reconstruct (CD_Interior_phi_Bh, DUGKS2DSolver.cpp:604-687), face
// macros (Update_MacroVar_h), BGK half step (Update_phi_h) and flux for the six
// populations with non-zero normal velocity.  bp: 18 planes of R1N, mac: 6 planes.
template <int DIR, int MODEL, int WALLS>
__device__ __forceinline__ void face_flux(const K32 &k, const float *bp, const float *mac, int ko, int kn, int step_t,
                                          float *hfl, float *ffl, int wall)
{
    constexpr bool HCZ = MODEL == MPD_MODEL_AC_HCZ;
    float hb[9], fb[9];
    const float hdt = k.hdt;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const int en = DIR == 0 ? ex9(q) : ey9(q);
        const int et = DIR == 0 ? ey9(q) : ex9(q);
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const float *b = bp + (s * 9 + q) * R1N;
            const float bo = b[ko], bn = b[kn];
            float v = 0.5f * (bo + bn);
            if (en != 0) v -= (hdt * k.ms * k.inv_dx * (float)en) * (bn - bo);
            if (et != 0)
                v -= (hdt * k.ms * k.inv_2dx * 0.5f * (float)et) *
                     ((b[kn + step_t] - b[kn - step_t]) + (b[ko + step_t] - b[ko - step_t]));
            if (s == 0) hb[q] = v; else fb[q] = v;
        }
    }
    const float ax = 0.5f * (mac[0 * R1N + ko] + mac[0 * R1N + kn]);
    const float ay = 0.5f * (mac[1 * R1N + ko] + mac[1 * R1N + kn]);
    const float rgx = 0.5f * (mac[2 * R1N + ko] + mac[2 * R1N + kn]);
    const float rgy = 0.5f * (mac[3 * R1N + ko] + mac[3 * R1N + kn]);
    const float Fx = 0.5f * (mac[4 * R1N + ko] + mac[4 * R1N + kn]);
    const float Fy = 0.5f * (mac[5 * R1N + ko] + mac[5 * R1N + kn]);
    float phi = 0.f;
#pragma unroll
    for (int q = 0; q < 9; ++q) phi += hb[q];
    const float rho = k.aphi * phi + k.bphi;
    float U = 0.f, V = 0.f, p, bh, ch;
    if (HCZ) {
        const float irr = rcp(rho * k.rt);
        const float sx = k.ms * ((fb[1] + fb[2] + fb[8]) - (fb[4] + fb[5] + fb[6]));
        const float sy = k.ms * ((fb[2] + fb[3] + fb[4]) - (fb[6] + fb[7] + fb[8]));
        U = (sx + 0.5f * hdt * Fx) * irr;
        V = (sy + 0.5f * hdt * Fy) * irr;
        float s = 0.f;
#pragma unroll
        for (int q = 0; q < 9; ++q) s += fb[q];
        p = s + 0.5f * hdt * k.rt * (rgx * U + rgy * V);
        const float tau = (k.nu_v + (phi - k.phi_v) * k.dnu) * k.inv_rt;
        bh = hdt * rcp(2.f * tau + hdt);
        ch = tau * bh;
    } else {
        float s = 0.f;
#pragma unroll
        for (int q = 1; q < 9; ++q) s += fb[q];
        p = k.kp * s;
        bh = k.bh0;
        ch = k.ch0;
    }
    const float uu = U * U + V * V;
    const float UF = U * Fx + V * Fy, UR = U * rgx + V * rgy;
    float ho[NZ], fo[NZ];
#pragma unroll
    for (int z = 0; z < NZ; ++z) {
        const int q = DIR == 0 ? (z == 0 ? 1 : z == 1 ? 2 : z == 2 ? 4 : z == 3 ? 5 : z == 4 ? 6 : 8)
                               : (z == 0 ? 2 : z == 1 ? 3 : z == 2 ? 4 : z == 3 ? 6 : z == 4 ? 7 : 8);
        const float ex = (float)ex9(q), ey = (float)ey9(q), w = k.w[q];
        float heq, feq, fso;
        if (HCZ) {
            const float u1 = k.ms_rt * (ex * U + ey * V);
            const float G = u1 * (1.f + 0.5f * u1) - uu * k.lambda0;
            heq = w * phi * (1.f + u1);
            feq = w * (p + rho * k.rt * G);
            const float xF = k.ms * (ex * Fx + ey * Fy) - UF;
            const float xR = k.ms * (ex * rgx + ey * rgy) - UR;
            fso = w * ((1.f + G) * xF + G * k.rt * xR);
        } else {
            heq = w * phi;
            feq = w * p * k.inv_rt;
            fso = w * k.ms * (ex * Fx + ey * Fy) * k.inv_rt;
        }
        const float hso = w * k.ms * (ex * ax + ey * ay);
        ho[z] = hb[q] + k.bh_h * (heq - hb[q]) + k.ch_h * hso;
        fo[z] = fb[q] + bh * (feq - fb[q]) + ch * fso;
    }
    if (WALLS && DIR == 1 && wall != 0) {
        // Wall_3_BB (Boundary.cpp:118-139): the y-face populations entering the
        // fluid (e_y > 0 at the bottom wall, < 0 at the top) take their partner's
        // value; order {2,3,4,6,7,8} pairs z with z+3
        float h0[NZ], f0[NZ];
#pragma unroll
        for (int z = 0; z < NZ; ++z) { h0[z] = ho[z]; f0[z] = fo[z]; }
#pragma unroll
        for (int z = 0; z < NZ; ++z) {
            const bool entering = wall < 0 ? (z < 3) : (z >= 3);
            if (entering) { ho[z] = h0[(z + 3) % NZ]; fo[z] = f0[(z + 3) % NZ]; }
        }
    }
#pragma unroll
    for (int z = 0; z < NZ; ++z) {
        const int q = DIR == 0 ? (z == 0 ? 1 : z == 1 ? 2 : z == 2 ? 4 : z == 3 ? 5 : z == 4 ? 6 : 8)
                               : (z == 0 ? 2 : z == 1 ? 3 : z == 2 ? 4 : z == 3 ? 6 : z == 4 ? 7 : 8);
        const float en = DIR == 0 ? (float)ex9(q) : (float)ey9(q);
        hfl[z] = ho[z] * (k.flx * en);
        ffl[z] = fo[z] * (k.flx * en);
    }
}

// One 16x15 tile of one case: stages A-D of the step.
template <int MODEL, int WALLS>
__device__ __forceinline__ void tile32(const K32 &k, int nx, int ny, const FusedPlanes32 &in, const FusedPlanes32 &out,
                                       int64_t N, int i0, int j0)
{
    constexpr bool HCZ = MODEL == MPD_MODEL_AC_HCZ;
    __shared__ float s_bp[18][R1N];        // bar-plus h(0..8) f(9..17); later y-face fluxes
    __shared__ float s_mac[6][R1N];        // AC source (2), grad rho (2), F (2)
    __shared__ float s_u[(2 * NZ + 1) * NXF];  // phi, phi_lag, then x-face fluxes (+ h mass flux)
    float *s_phi = s_u, *s_phl = s_u + R2N;
    static_assert(2 * R2N <= 2 * NZ * NXF, "union too small");
    static_assert((2 * NZ + 1) * NYF <= 18 * R1N, "y-face fluxes do not fit");
    const int tid = threadIdx.x;

    // ---- stage A ----
    for (int q = tid; q < R2N; q += NT) {
        int gi = wrapc(i0 - 2 + q % R2X, nx), gj = j0 - 2 + q / R2X;
        float p = 0.f, pl = 0.f;
        if (WALLS && (gj < 0 || gj >= ny)) {
            if (gj == -1 || gj == ny) {
                int jn = gj < 0 ? 0 : ny - 1, jnn = gj < 0 ? 1 : ny - 2;
                p = 2.f * in.phl[(size_t)jn * nx + gi] - in.phl[(size_t)jnn * nx + gi];
                if (HCZ) pl = in.gl[(gj < 0 ? 0 : nx) + gi];
            }
        } else {
            gj = wrapc(gj, ny);
            p = in.phi[(size_t)gj * nx + gi];
            if (HCZ) pl = in.phl[(size_t)gj * nx + gi];
        }
        s_phi[q] = p;
        if (HCZ) s_phl[q] = pl;
    }
    __syncthreads();

    // ---- stage B: macros and bar-plus ----
    for (int c = tid; c < R1N; c += NT) {
        const int rx = c % R1X, ry = c / R1X;
        const int gi = wrapc(i0 - 1 + rx, nx);
        int gj = j0 - 1 + ry;
        if (WALLS && (gj < 0 || gj >= ny)) continue;
        gj = wrapc(gj, ny);
        const size_t n = (size_t)gj * nx + gi;
        float h[9], f[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            h[q] = in.h[(size_t)q * N + n];
            f[q] = in.f[(size_t)q * N + n];
        }
        const int c2 = (ry + 1) * R2X + (rx + 1);
        const float phi = s_phi[c2];
        const float pc[4] = {s_phi[c2 + 1], s_phi[c2 + R2X], s_phi[c2 - 1], s_phi[c2 - R2X]};
        const float pd[4] = {s_phi[c2 + R2X + 1], s_phi[c2 + R2X - 1], s_phi[c2 - R2X - 1], s_phi[c2 - R2X + 1]};
        const float rho = k.aphi * phi + k.bphi;
        float U = 0.f, V = 0.f, p, bbp, cbp;
        if (HCZ) {
            // Update_MacroVar with the previous step's sources (D2Q9HCZ.cpp:133-182)
            const float lc[4] = {s_phl[c2 + 1], s_phl[c2 + R2X], s_phl[c2 - 1], s_phl[c2 - R2X]};
            const float ld[4] = {s_phl[c2 + R2X + 1], s_phl[c2 + R2X - 1], s_phl[c2 - R2X - 1], s_phl[c2 - R2X + 1]};
            float lgx, lgy, llap, lFx, lFy, lrx, lry;
            stencil(k, s_phl[c2], lc, ld, lgx, lgy, llap);
            momentum_source(k, s_phl[c2], lgx, lgy, llap, lFx, lFy, lrx, lry);
            const float irr = rcp(rho * k.rt);
            const float sx = k.ms * ((f[1] + f[2] + f[8]) - (f[4] + f[5] + f[6]));
            const float sy = k.ms * ((f[2] + f[3] + f[4]) - (f[6] + f[7] + f[8]));
            U = (sx + k.hdt * lFx) * irr;
            V = (sy + k.hdt * lFy) * irr;
            float s = 0.f;
#pragma unroll
            for (int q = 0; q < 9; ++q) s += f[q];
            p = s + k.hdt * k.rt * (lrx * U + lry * V);
            const float tau = (k.nu_v + (phi - k.phi_v) * k.dnu) * k.inv_rt;
            bbp = 1.5f * k.dt * rcp(2.f * tau + k.dt);   // 1 - (2 tau - dt/2)/(2 tau + dt)
            cbp = tau * bbp;
        } else {
            // AC-p: U = V = 0, p = Kp * sum_{q != 0} f (D2Q9AC.cpp:197-203)
            float s = 0.f;
#pragma unroll
            for (int q = 1; q < 9; ++q) s += f[q];
            p = k.kp * s;
            bbp = k.bbp0;
            cbp = k.cbp0;
        }
        // MacroSource (SourceM.cpp:35-62) from the current phi stencil
        float gx, gy, lap, Fx, Fy, rgx, rgy;
        stencil(k, phi, pc, pd, gx, gy, lap);
        momentum_source(k, phi, gx, gy, lap, Fx, Fy, rgx, rgy);
        float ax, ay, npu = 0.f, npv = 0.f;
        {
            const float L2 = gx * gx + gy * gy;
            float dpu = 0.f, dpv = 0.f;
            if (HCZ) {
                npu = phi * U;
                npv = phi * V;
                dpu = (npu - in.ppu[n]) * k.inv_rtdt;
                dpv = (npv - in.ppv[n]) * k.inv_rtdt;
            }
            if (L2 != 0.f) {
                const float t = -4.f * (phi - k.phi_l) * (phi - k.phi_v) * k.inv_wdphi * rsqrtf(L2);
                ax = gx * t + dpu;
                ay = gy * t + dpv;
            } else {
                ax = dpu;
                ay = dpv;
            }
        }
        // Update_phi_Eq / Update_DVDF_Source / Update_phi_BP (D2Q9HCZ.cpp:54-92,244-257;
        // D2Q9AC.cpp:69-105; DUGKS2DSolver.cpp:379-401)
        const float uu = U * U + V * V;
        const float UF = U * Fx + V * Fy, UR = U * rgx + V * rgy;
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            const float ex = (float)ex9(q), ey = (float)ey9(q), w = k.w[q];
            const float eu = ex * U + ey * V;            // e . u (compile-time e)
            const float eA = k.ms * (ex * ax + ey * ay);  // xi . AC source
            float heq, feq, fso;
            if (HCZ) {
                const float u1 = k.ms_rt * eu;
                const float G = u1 * (1.f + 0.5f * u1) - uu * k.lambda0;
                heq = w * (1.f + u1) * phi;
                feq = w * (p + rho * k.rt * G);
                const float xF = k.ms * (ex * Fx + ey * Fy) - UF;
                const float xR = k.ms * (ex * rgx + ey * rgy) - UR;
                fso = w * ((1.f + G) * xF + G * k.rt * xR);
            } else {
                heq = w * phi;
                feq = (w - (q == 0 ? 1.f : 0.f)) * p * k.inv_rt;
                fso = w * k.ms * (ex * Fx + ey * Fy) * k.inv_rt;
            }
            const float hso = w * eA;
            s_bp[q][c] = h[q] + k.bbp_h * (heq - h[q]) + k.cbp_h * hso;
            s_bp[9 + q][c] = f[q] + bbp * (feq - f[q]) + cbp * fso;
        }
        s_mac[0][c] = ax; s_mac[1][c] = ay; s_mac[2][c] = rgx;
        s_mac[3][c] = rgy; s_mac[4][c] = Fx; s_mac[5][c] = Fy;
        const bool interior = rx >= 1 && rx <= TX && ry >= 1 && ry <= TY && i0 - 1 + rx < nx && j0 - 1 + ry < ny;
        if (interior && (HCZ || WALLS)) out.phl[n] = phi;
        if (interior && HCZ) {
            out.ppu[n] = npu;
            out.ppv[n] = npv;
            if (WALLS && (gj == 0 || gj == ny - 1)) {
                const int nxt = gj == 0 ? c2 + R2X : c2 - R2X;
                out.gl[(gj == 0 ? 0 : nx) + gi] = 2.f * s_phl[c2] - s_phl[nxt];
            }
        }
    }
    __syncthreads();

    if (WALLS) {
        for (int c = tid; c < R1N; c += NT) {
            const int rx = c % R1X, ry = c / R1X;
            const int gj = j0 - 1 + ry;
            if (gj != -1 && gj != ny) continue;
            const int dn = gj < 0 ? 1 : -1;
            const int kn = (ry + dn) * R1X + rx, knn = (ry + 2 * dn) * R1X + rx;
#pragma unroll
            for (int q = 0; q < 18; ++q) s_bp[q][c] = 2.f * s_bp[q][kn] - s_bp[q][knn];
#pragma unroll
            for (int a = 0; a < 6; ++a) s_mac[a][c] = 2.f * s_mac[a][kn] - s_mac[a][knn];
        }
        __syncthreads();
    }

    // ---- stage C: faces ----
    if (tid < NXF) {
        const int fi = tid % (TX + 1), fj = tid / (TX + 1);
        const int ko = (fj + 1) * R1X + fi;
        float hfl[NZ], ffl[NZ];
        face_flux<0, MODEL, WALLS>(k, &s_bp[0][0], &s_mac[0][0], ko, ko + 1, R1X, hfl, ffl, 0);
        float mass = 0.f;
#pragma unroll
        for (int z = 0; z < NZ; ++z) {
            s_u[z * NXF + tid] = hfl[z];
            s_u[(NZ + z) * NXF + tid] = ffl[z];
            mass += hfl[z];
        }
        s_u[2 * NZ * NXF + tid] = mass;
    }
    float yh[NZ], yf[NZ];
    {
        const int fi = tid % TX, fj = tid / TX;
        const int ko = fj * R1X + fi + 1;
        const int gj = j0 + fj;
        const int wall = WALLS ? (gj == 0 ? -1 : (gj == ny ? 1 : 0)) : 0;
        face_flux<1, MODEL, WALLS>(k, &s_bp[0][0], &s_mac[0][0], ko, ko + R1X, 1, yh, yf, wall);
    }
    const int ci = tid % TX, cj = tid / TX;
    const bool cell = tid < TX * TY;
    const int kc = (cj + 1) * R1X + ci + 1;
    float hbp[9], fbp[9];
    if (cell) {
#pragma unroll
        for (int q = 0; q < 9; ++q) { hbp[q] = s_bp[q][kc]; fbp[q] = s_bp[9 + q][kc]; }
    }
    __syncthreads();
    {
        float mass = 0.f;
#pragma unroll
        for (int z = 0; z < NZ; ++z) {
            s_bp[z][tid] = yh[z];
            s_bp[NZ + z][tid] = yf[z];
            mass += yh[z];
        }
        s_bp[2 * NZ][tid] = mass;
    }
    __syncthreads();

    // ---- stage D: Update_phi_T as b + (b - x~)/3 + dt/V * (W - E + S - N) ----
    const int gi = i0 + ci, gj = j0 + cj;
    if (!cell || gi >= nx || gj >= ny) return;
    const size_t n = (size_t)gj * nx + gi;
    const int kW = cj * (TX + 1) + ci, kE = kW + 1;
    const int kS = cj * TX + ci, kN = kS + TX;
    float fluxh[9], fluxf[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) { fluxh[q] = 0.f; fluxf[q] = 0.f; }
#pragma unroll
    for (int z = 0; z < NZ; ++z) {
        const int qx = xq(z), qy = yq(z);
        fluxh[qx] += s_u[z * NXF + kW] - s_u[z * NXF + kE];
        fluxf[qx] += s_u[(NZ + z) * NXF + kW] - s_u[(NZ + z) * NXF + kE];
        fluxh[qy] += s_bp[z][kS] - s_bp[z][kN];
        fluxf[qy] += s_bp[NZ + z][kS] - s_bp[NZ + z][kN];
    }
    // phi is the conserved variable: it moves only by the antisymmetric face
    // mass fluxes (each face's sum over populations, shared by both cells), and
    // h~_0 takes the remainder so that sum_q h~ = phi.  In exact arithmetic this
    // is the same update; in FP32 it keeps the rounding of the per-population
    // products from drifting the phase-field mass.
    const float third = 1.f / 3.f;
    const float mW = s_u[2 * NZ * NXF + kW], mE = s_u[2 * NZ * NXF + kE];
    const float mS = s_bp[2 * NZ][kS], mN = s_bp[2 * NZ][kN];
    const float phi = in.phi[n] + k.dtv * ((mW - mE) + (mS - mN));
    float rest = 0.f;
#pragma unroll
    for (int q = 1; q < 9; ++q) {
        const float hn = hbp[q] + third * (hbp[q] - in.h[(size_t)q * N + n]) + k.dtv * fluxh[q];
        out.h[(size_t)q * N + n] = hn;
        rest += hn;
    }
    out.h[n] = phi - rest;
#pragma unroll
    for (int q = 0; q < 9; ++q)
        out.f[(size_t)q * N + n] = fbp[q] + third * (fbp[q] - in.f[(size_t)q * N + n]) + k.dtv * fluxf[q];
    out.phi[n] = phi;
}

template <int MODEL, int WALLS>
__global__ void __launch_bounds__(NT) k_fused32(K32 k, int nx, int ny, FusedPlanes32 in, FusedPlanes32 out, int64_t N,
                                               int ntx, int ntiles)
{
    const int tile = tile_of_block(blockIdx.x, ntiles);
    if (tile < 0) return;
    tile32<MODEL, WALLS>(k, nx, ny, in, out, N, (tile % ntx) * TX, (tile / ntx) * TY);
}

// Batched step: every workgroup takes one (case, tile) entry of a work list
// built on the host; the cases share the model and the y boundary.
template <int MODEL, int WALLS>
__global__ void __launch_bounds__(NT) k_fused32_batch(const BatchCase *cases, const int2 *work, int nwork, int parity)
{
    const int b = blockIdx.x;
    if (b >= nwork) return;
    const int2 w = work[b];
    if (w.x < 0) return;
    const BatchCase &c = cases[w.x];
    const int sel = c.cur ^ parity;
    const int ntx = (c.nx + TX - 1) / TX;
    tile32<MODEL, WALLS>(c.k, c.nx, c.ny, c.pl[sel], c.pl[sel ^ 1], c.N, (w.y % ntx) * TX, (w.y / ntx) * TY);
}

K32 make_k32(const Consts &C)
{
    K32 k;
    const double ms = C.xi[0][1];
    k.ms = (float)ms;
    k.ms_rt = (float)(ms / C.rt);
    k.rt = (float)C.rt;
    k.inv_rt = (float)(1.0 / C.rt);
    k.lambda0 = (float)C.lambda0;
    k.hdt = (float)C.hdt;
    k.dt = (float)C.dt;
    k.aphi = (float)C.aphi;
    k.bphi = (float)C.bphi;
    k.beta = (float)C.beta;
    k.kappa = (float)C.kappa;
    k.phi_l = (float)C.phi_l;
    k.phi_v = (float)C.phi_v;
    k.nu_v = (float)C.nu_v;
    k.dnu = (float)(C.nu_l - C.nu_v);
    k.inv_d2x6 = (float)(1.0 / C.d2x6);
    k.inv_lapden = (float)(1.0 / C.lapden);
    k.inv_rtdt = (float)(1.0 / C.rtdt);
    k.inv_wdphi = (float)(1.0 / C.wdphi);
    k.inv_dx = (float)(1.0 / C.dx);
    k.inv_2dx = (float)(1.0 / C.twodx);
    k.kp = (float)C.kp;
    k.dtv = (float)C.dtv;
    k.gx = (float)C.gx;
    k.gy = (float)C.gy;
    k.rho_ref = (float)C.rho_ref;
    k.bbp_h = (float)(1.5 * C.dt / (2.0 * C.taum + C.dt));
    k.cbp_h = (float)(C.taum * 1.5 * C.dt / (2.0 * C.taum + C.dt));
    k.bh_h = (float)(C.hdt / (2.0 * C.taum + C.hdt));
    k.ch_h = (float)(C.taum * C.hdt / (2.0 * C.taum + C.hdt));
    k.tau0 = (float)C.tau0;
    // AC-p: tau = Mu/(rho RT) with Mu = rho RT Tau0 (D2Q9AC.cpp:34-60) -> Tau0
    k.bbp0 = (float)(1.5 * C.dt / (2.0 * C.tau0 + C.dt));
    k.cbp0 = (float)(C.tau0 * 1.5 * C.dt / (2.0 * C.tau0 + C.dt));
    k.bh0 = (float)(C.hdt / (2.0 * C.tau0 + C.hdt));
    k.ch0 = (float)(C.tau0 * C.hdt / (2.0 * C.tau0 + C.hdt));
    k.flx = (float)(ms * C.dx);
    // weights rounded to FP32 with w0 = 1 - 4 w1 - 4 w2 so that they sum to 1 exactly
    const float w1 = 1.f / 9.f, w2 = 1.f / 36.f;
    const float w0 = 1.f - 4.f * w1 - 4.f * w2;
    for (int q = 0; q < 9; ++q) k.w[q] = q == 0 ? w0 : (q % 2 ? w1 : w2);
    return k;
}

}  // namespace

struct Batch32Plan {
    BatchCase *d_cases = nullptr;
    int2 *d_work = nullptr;
    int ncases = 0, nwork = 0, model = 0, walls = 0;
    size_t cap_cases = 0, cap_work = 0;
};

Batch32Plan *batch32_new() { return new Batch32Plan(); }

void batch32_free(Batch32Plan *P)
{
    if (!P) return;
    if (P->d_cases) (void)hipFree(P->d_cases);
    if (P->d_work) (void)hipFree(P->d_work);
    delete P;
}

// (Re)build the device tables: cases in order, tiles row-major per case, then
// dealt so that XCD b%8 receives a contiguous run of the list.
int batch32_plan(Batch32Plan *P, const BatchMember *m, int n, hipStream_t st)
{
    std::vector<BatchCase> cs((size_t)n);
    std::vector<int2> list;
    for (int c = 0; c < n; ++c) {
        const Consts &C = *m[c].C;
        BatchCase &b = cs[(size_t)c];
        b.k = make_k32(C);
        b.nx = C.nx;
        b.ny = C.ny;
        b.N = m[c].N;
        b.pl[0] = m[c].pl[0];
        b.pl[1] = m[c].pl[1];
        b.cur = m[c].cur;
        const int ntiles = ((C.nx + TX - 1) / TX) * ((C.ny + TY - 1) / TY);
        for (int t = 0; t < ntiles; ++t) list.push_back(make_int2(c, t));
    }
    const int per = ((int)list.size() + 7) / 8;
    std::vector<int2> work((size_t)per * 8, make_int2(-1, 0));
    for (int b = 0; b < per * 8; ++b) {
        const int p = (b % 8) * per + b / 8;
        if (p < (int)list.size()) work[(size_t)b] = list[(size_t)p];
    }
    if ((size_t)n > P->cap_cases) {
        if (P->d_cases) (void)hipFree(P->d_cases);
        if (hipMalloc(&P->d_cases, sizeof(BatchCase) * (size_t)n) != hipSuccess) return -1;
        P->cap_cases = (size_t)n;
    }
    if (work.size() > P->cap_work) {
        if (P->d_work) (void)hipFree(P->d_work);
        if (hipMalloc(&P->d_work, sizeof(int2) * work.size()) != hipSuccess) return -1;
        P->cap_work = work.size();
    }
    if (n && hipMemcpyAsync(P->d_cases, cs.data(), sizeof(BatchCase) * (size_t)n, hipMemcpyHostToDevice, st) != hipSuccess)
        return -1;
    if (!work.empty() &&
        hipMemcpyAsync(P->d_work, work.data(), sizeof(int2) * work.size(), hipMemcpyHostToDevice, st) != hipSuccess)
        return -1;
    if (hipStreamSynchronize(st) != hipSuccess) return -1;
    P->ncases = n;
    P->nwork = (int)work.size();
    P->model = n ? m[0].C->model : 0;
    P->walls = n ? m[0].C->walls : 0;
    return 0;
}

void batch32_step(const Batch32Plan *P, int parity, hipStream_t st)
{
    if (!P->nwork) return;
    dim3 g((unsigned)P->nwork), b(NT);
    if (P->model == MPD_MODEL_AC_HCZ) {
        if (P->walls) hipLaunchKernelGGL((k_fused32_batch<MPD_MODEL_AC_HCZ, 1>), g, b, 0, st, P->d_cases, P->d_work, P->nwork, parity);
        else hipLaunchKernelGGL((k_fused32_batch<MPD_MODEL_AC_HCZ, 0>), g, b, 0, st, P->d_cases, P->d_work, P->nwork, parity);
    } else {
        if (P->walls) hipLaunchKernelGGL((k_fused32_batch<MPD_MODEL_AC_P, 1>), g, b, 0, st, P->d_cases, P->d_work, P->nwork, parity);
        else hipLaunchKernelGGL((k_fused32_batch<MPD_MODEL_AC_P, 0>), g, b, 0, st, P->d_cases, P->d_work, P->nwork, parity);
    }
}

void fused32_step(const FusedPlanes32 &in, const FusedPlanes32 &out, const Consts &C, int64_t N, hipStream_t st)
{
    const K32 k = make_k32(C);
    const int ntx = (C.nx + TX - 1) / TX, nty = (C.ny + TY - 1) / TY;
    const int ntiles = ntx * nty;
    const int nblocks = ((ntiles + 7) / 8) * 8;
    if (C.model == MPD_MODEL_AC_HCZ) {
        if (C.walls) hipLaunchKernelGGL((k_fused32<MPD_MODEL_AC_HCZ, 1>), dim3(nblocks), dim3(NT), 0, st, k, C.nx, C.ny, in, out, N, ntx, ntiles);
        else hipLaunchKernelGGL((k_fused32<MPD_MODEL_AC_HCZ, 0>), dim3(nblocks), dim3(NT), 0, st, k, C.nx, C.ny, in, out, N, ntx, ntiles);
    } else {
        if (C.walls) hipLaunchKernelGGL((k_fused32<MPD_MODEL_AC_P, 1>), dim3(nblocks), dim3(NT), 0, st, k, C.nx, C.ny, in, out, N, ntx, ntiles);
        else hipLaunchKernelGGL((k_fused32<MPD_MODEL_AC_P, 0>), dim3(nblocks), dim3(NT), 0, st, k, C.nx, C.ny, in, out, N, ntx, ntiles);
    }
}

}  // namespace mpd
