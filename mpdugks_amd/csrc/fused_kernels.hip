// Fused DUGKS step: one kernel launch per time step for the Allen-Cahn models
// (AC-p, AC+HCZ) with the CD face scheme, periodic in x; y periodic or walls
// (bounce-back or non-equilibrium extrapolation).
//
// Derived state.  Between steps only what the reference carries and cannot
// recompute is kept, in structure-of-arrays planes (ping-pong):
//   h~[9], f~[9]          populations
//   phi                   sum of h~ (the cell's Phi after Update_MacroVar)
//   phi_lag               Phi of the previous state (HCZ): the momentum
//                         sources F, grad(rho) that Update_MacroVar used were
//                         computed from it by the previous MacroSource
//                         (Ark/Cpp/DUGKS2DSolver.cpp:284-295, SourceM.cpp:27-34)
//   ppu, ppv              prevPhiU/V before the last MacroSource (HCZ)
//   gl[2][nx]             walls: ghost Phi that went with phi_lag's stencil
// Every cell macro of the reference (U, V, p, tau, F, grad rho, the AC source
// vector ...) is rebuilt inside the step from these planes with the same
// expressions, so the FP64 kernel reproduces the phase path bit for bit.
//
// FP32-bulk: planes are float; f~ is stored as the deviation from the
// equilibrium of the reference pressure p_ref at rest (w_q p_ref for HCZ,
// (w_q - delta_q0) p_ref/RT for AC-p).  Every expression that touches p is
// linear in p, so the same code runs with p read as p - p_ref.
//
// Tile: 16 x 15 or 16 x 16 cells per 256-thread workgroup.  LDS holds
// bar-plus (18 planes) and the six face-averaged sources on the tile plus a
// one-cell halo, phi/phi_lag on a two-cell halo, and the face fluxes.  16 x 15:
// x-faces (17x15 = 255) and y-faces (16x16 = 256) map one face per thread;
// 16 x 16: 272 of each, the last 16 of each direction on one wave each in a
// second pass.
#include <hip/hip_runtime.h>

#include "fused.h"
#include "physics.h"

namespace mpd {

namespace {

constexpr int NTHREADS = 256;

template <int TX_, int TY_>
struct Tile64 {
    static constexpr int TX = TX_, TY = TY_;
    static constexpr int R1X = TX + 2, R1Y = TY + 2;  // one-cell halo (bar-plus, sources)
    static constexpr int R2X = TX + 4, R2Y = TY + 4;  // two-cell halo (phi stencils)
    static constexpr int R1N = R1X * R1Y, R2N = R2X * R2Y;
    static constexpr int NXF = (TX + 1) * TY;         // x-faces
    static constexpr int NYF = TX * (TY + 1);         // y-faces
    // faces beyond one per thread: a second pass on wave 2 (x) and wave 3 (y)
    static constexpr int EXF = NXF > NTHREADS ? NXF - NTHREADS : 0;
    static constexpr int EYF = NYF > NTHREADS ? NYF - NTHREADS : 0;
    static_assert(TX * TY <= NTHREADS && EXF <= 64 && EYF <= 64, "one cell and at most two face passes per thread");
    static_assert(12 * (NXF + NYF) <= (2 * Q9 + 6) * R1N, "face fluxes must fit over the bar-plus planes");
};

template <typename T>
struct Planes {
    T *h, *f, *phi, *phl, *ppu, *ppv, *gl;
    double *phd;  // FP32-bulk: the conserved phi in FP64 (nullptr for the FP64 state)
};

// periodic index for any offset (small meshes may wrap more than once per tile)
__device__ inline int wrapc(int i, int n)
{
    int r = i % n;
    return r < 0 ? r + n : r;
}

// tile index with XCD-aware ordering: workgroups are dealt round-robin to the
// 8 XCDs, so give XCD k a contiguous band of tiles (vertical neighbours share L2)
__device__ inline int tile_of_block(int b, int ntiles)
{
    const int nxcd = 8;
    int per = (ntiles + nxcd - 1) / nxcd;
    int xcd = b % nxcd, loc = b / nxcd;
    int t = xcd * per + loc;
    return t < ntiles ? t : -1;
}

template <typename T>
__device__ inline void grad_lap(const Consts &C, const T *c, const T *d, T phi, T &gx, T &gy, T &lap)
{
    grad9<T>(C, c, d, gx, gy);
    lap = ((T)4 * (c[0] + c[1] + c[2] + c[3]) + (d[0] + d[1] + d[2] + d[3]) - (T)20 * phi) / (T)C.lapden;
    set_zero(gx);
    set_zero(gy);
}

// F and grad(rho) that MacroSource produced from the phi stencil (c, d)
// (SourceMomentum, Ark/Cpp/SourceM.cpp:19-34)
template <typename T>
__device__ inline void lag_sources(const Consts &C, T phi, const T *c, const T *d, T &Fx, T &Fy, T &rx, T &ry)
{
    T gx, gy, lap;
    grad_lap<T>(C, c, d, phi, gx, gy, lap);
    T mu_c = (T)2 * (T)C.beta * (phi - (T)C.phi_l) * (phi - (T)C.phi_v) * ((T)2 * phi - (T)C.phi_l - (T)C.phi_v) -
             (T)C.kappa * lap;
    rx = (T)C.aphi * gx;
    ry = (T)C.aphi * gy;
    Fx = mu_c * gx;
    Fy = mu_c * gy;
    T rho = (T)C.aphi * phi + (T)C.bphi;
    if (C.gx != 0.0) Fx -= (T)C.gx * (rho - (T)C.rho_ref);
    if (C.gy != 0.0) Fy -= (T)C.gy * (rho - (T)C.rho_ref);
}

// All reference macros of one cell from the derived state.  pc/pd: phi
// stencil (E,N,W,S / NE,NW,SW,SE), lc/ld: phi_lag stencil (HCZ).
template <typename T, int MODEL, bool UNIT = false>
__device__ inline void derive_cell(const Consts &C, const T *ht, const T *ft, T phi_lag, const T *lc, const T *ld,
                                   T ppu, T ppv, const T *pc, const T *pd, CellMacro<T> &m, AcSource<T> &s, T &mu)
{
    T Fl_x = 0, Fl_y = 0, rl_x = 0, rl_y = 0;
    if (MODEL == MPD_MODEL_AC_HCZ) lag_sources<T>(C, phi_lag, lc, ld, Fl_x, Fl_y, rl_x, rl_y);
    cell_macrovar_ac<T, MODEL, UNIT>(C, ht, ft, Fl_x, Fl_y, rl_x, rl_y, m.phi, m.rho, m.U, m.V, m.p, mu, m.tau);
    ac_source<T>(C, m.phi, m.rho, m.U, m.V, ppu, ppv, pc, pd, s);
    m.px = s.sx;
    m.py = s.sy;
    m.rx = s.rx;
    m.ry = s.ry;
    m.Fx = s.Fx;
    m.Fy = s.Fy;
}

// x-face t of a tile: face (fi, fj) between R1 cells (fi, fj+1) and
// (fi+1, fj+1); fluxes of the six populations with e_x != 0
template <typename T, int MODEL, bool UNIT, class G>
__device__ inline void xface64(const Consts &C, const T (*s_bp)[G::R1N], const T (*s_mac)[G::R1N], int t, T *xh,
                               T *xf)
{
    constexpr int TX = G::TX, R1X = G::R1X;
    constexpr int MX[6] = {1, 2, 4, 5, 6, 8};
    const T hdt = (T)C.hdt;
    const int fi = t % (TX + 1), fj = t / (TX + 1);
    const int ko = (fj + 1) * R1X + fi, kn = ko + 1;
    T hb[Q9], fb[Q9];
#pragma unroll
    for (int q = 0; q < 2 * Q9; ++q) {
        const T *b = s_bp[q];
        T v = cd_face_value<T, UNIT>(C, 0, q % Q9, b[ko], b[kn], b[kn + R1X], b[kn - R1X], b[ko + R1X], b[ko - R1X]);
        if (q < Q9) hb[q] = v; else fb[q - Q9] = v;
    }
    FaceMacro<T> fm;
    fm.px = (T)0.5 * (s_mac[0][ko] + s_mac[0][kn]);
    fm.py = (T)0.5 * (s_mac[1][ko] + s_mac[1][kn]);
    fm.rx = (T)0.5 * (s_mac[2][ko] + s_mac[2][kn]);
    fm.ry = (T)0.5 * (s_mac[3][ko] + s_mac[3][kn]);
    fm.Fx = (T)0.5 * (s_mac[4][ko] + s_mac[4][kn]);
    fm.Fy = (T)0.5 * (s_mac[5][ko] + s_mac[5][kn]);
    face_macros_ac<T, MODEL, UNIT>(C, hb, fb, fm);
    T ah = (T)2.0 * fm.tau / ((T)2.0 * fm.tau + hdt);
    T bh = (T)1.0 - ah;
    T ch = fm.tau * bh;
#pragma unroll
    for (int z = 0; z < 6; ++z) {
        const int q = MX[z];
        T hh, ff;
        face_bgk<T, MODEL, UNIT>(C, fm, q, hb[q], fb[q], ah, bh, ch, hh, ff);
        const T xin = (T)C.dx * xi_of<UNIT, T>(C, 0, q);
        xh[z] = hh * xin;
        xf[z] = ff * xin;
    }
}

// y-face t of a tile: face (fi, fj) between R1 cells (fi+1, fj) and
// (fi+1, fj+1); fluxes of the six populations with e_y != 0, wall faces
// included
template <typename T, int MODEL, int WALLS, bool UNIT, class G>
__device__ inline void yface64(const Consts &C, const T (*s_bp)[G::R1N], const T (*s_mac)[G::R1N], const T *s_nee,
                               int t, int j0, int ny, T *yh, T *yf)
{
    constexpr int TX = G::TX, R1X = G::R1X;
    constexpr int MY[6] = {2, 3, 4, 6, 7, 8};
    const T hdt = (T)C.hdt;
    const int fi = t % TX, fj = t / TX;
    const int ko = fj * R1X + fi + 1, kn = ko + R1X;
    T hb[Q9], fb[Q9];
#pragma unroll
    for (int q = 0; q < 2 * Q9; ++q) {
        const T *b = s_bp[q];
        T v = cd_face_value<T, UNIT>(C, 1, q % Q9, b[ko], b[kn], b[kn + 1], b[kn - 1], b[ko + 1], b[ko - 1]);
        if (q < Q9) hb[q] = v; else fb[q - Q9] = v;
    }
    FaceMacro<T> fm;
    fm.px = (T)0.5 * (s_mac[0][ko] + s_mac[0][kn]);
    fm.py = (T)0.5 * (s_mac[1][ko] + s_mac[1][kn]);
    fm.rx = (T)0.5 * (s_mac[2][ko] + s_mac[2][kn]);
    fm.ry = (T)0.5 * (s_mac[3][ko] + s_mac[3][kn]);
    fm.Fx = (T)0.5 * (s_mac[4][ko] + s_mac[4][kn]);
    fm.Fy = (T)0.5 * (s_mac[5][ko] + s_mac[5][kn]);
    face_macros_ac<T, MODEL, UNIT>(C, hb, fb, fm);
    T ah = (T)2.0 * fm.tau / ((T)2.0 * fm.tau + hdt);
    T bh = (T)1.0 - ah;
    T ch = fm.tau * bh;
    // face values of the populations with e_y != 0 (wall faces: the
    // incoming ones are replaced below by their partners, also e_y != 0)
    T vh[Q9], vf[Q9];
#pragma unroll
    for (int z = 0; z < 6; ++z) {
        const int q = MY[z];
        face_bgk<T, MODEL, UNIT>(C, fm, q, hb[q], fb[q], ah, bh, ch, vh[q], vf[q]);
    }
    const int gj = j0 + fj;
    if (WALLS == 2 && (gj == 0 || gj == ny)) {
        // Wall_3_NEE: the wall-row cell's values (stage B), then bounce-back
        const T *v = s_nee + ((gj == 0 ? 0 : 1) * TX + fi) * 2 * Q9;
#pragma unroll
        for (int z = 0; z < 6; ++z) { vh[MY[z]] = v[MY[z]]; vf[MY[z]] = v[Q9 + MY[z]]; }
    }
    if (WALLS && (gj == 0 || gj == ny)) {
        // Wall_3_BB (Ark/Cpp/Boundary.cpp:118-139): incoming take the partner's value
        T h0[Q9], f0[Q9];
#pragma unroll
        for (int z = 0; z < 6; ++z) { h0[MY[z]] = vh[MY[z]]; f0[MY[z]] = vf[MY[z]]; }
#pragma unroll
        for (int z = 0; z < 6; ++z) {
            const int q = MY[z];
            const bool inc = gj == 0 ? (ey9(q) > 0) : (ey9(q) < 0);
            if (inc) { vh[q] = h0[bb9(q)]; vf[q] = f0[bb9(q)]; }
        }
    }
#pragma unroll
    for (int z = 0; z < 6; ++z) {
        const int q = MY[z];
        const T xin = (T)C.dx * xi_of<UNIT, T>(C, 1, q);
        yh[z] = vh[q] * xin;
        yf[z] = vf[q] * xin;
    }
}

// Two workgroups per CU (the LDS allows two): the generic (non-UNIT) AC-p
// instantiations would take a few registers over 256 and fall to one; with the
// bound they spill a few dwords instead.
template <typename T, int MODEL, int WALLS, bool UNIT, class G>
__global__ void __launch_bounds__(NTHREADS, 2) k_fused_step(Consts C, Planes<T> in, Planes<T> out, int64_t N, int ntx,
                                                         int ntiles)
{
    constexpr bool HCZ = MODEL == MPD_MODEL_AC_HCZ;
    constexpr int TX = G::TX, TY = G::TY, R1X = G::R1X, R2X = G::R2X, R1N = G::R1N, R2N = G::R2N;
    constexpr int NXF = G::NXF, NYF = G::NYF;
    // bar-plus (h in 0..8, f in 9..17) and the six face sources px, py, rx,
    // ry, Fx, Fy on the one-cell halo; after the faces, the face fluxes of the
    // six populations with a non-zero normal velocity (x: 12 x NXF, y: 12 x NYF)
    __shared__ T s_big[(2 * Q9 + 6) * R1N];
    __shared__ T s_ph[2 * R2N];         // phi, phi_lag on the two-cell halo
    __shared__ T s_nee[WALLS == 2 ? 2 * TX * 2 * Q9 : 1];  // NEE wall faces: [bottom, top][column][h, f]
    T(*s_bp)[R1N] = reinterpret_cast<T(*)[R1N]>(s_big);
    T(*s_mac)[R1N] = reinterpret_cast<T(*)[R1N]>(s_big + 2 * Q9 * R1N);
    T *s_phi = s_ph, *s_phl = s_ph + R2N;
    T *s_fx = s_big, *s_fy = s_big + 12 * NXF;

    const int tile = tile_of_block(blockIdx.x, ntiles);
    if (tile < 0) return;
    const int nx = C.nx, ny = C.ny;
    const int i0 = (tile % ntx) * TX, j0 = (tile / ntx) * TY;
    const int tid = threadIdx.x;

    // ---- stage A: phi and phi_lag on the two-cell halo ----
    for (int k = tid; k < R2N; k += NTHREADS) {
        int gi = wrapc(i0 - 2 + k % R2X, nx), gj = j0 - 2 + k / R2X;
        T p = 0, pl = 0;
        if (WALLS && (gj < 0 || gj >= ny)) {
            if (gj == -1 || gj == ny) {
                // WallShadowC_fBP ghost of the current phi: 2*near - next of phi_lag,
                // the Phi the reference extrapolated at the start of the last step
                int jn = gj < 0 ? 0 : ny - 1, jnn = gj < 0 ? 1 : ny - 2;
                if (HCZ) {
                    p = (T)2.0 * in.phl[(size_t)jn * nx + gi] - in.phl[(size_t)jnn * nx + gi];
                    pl = in.gl[(gj < 0 ? 0 : nx) + gi];
                } else {
                    p = (T)2.0 * in.phl[(size_t)jn * nx + gi] - in.phl[(size_t)jnn * nx + gi];
                }
            }
        } else {
            gj = wrapc(gj, ny);
            p = in.phi[(size_t)gj * nx + gi];
            if (HCZ) pl = in.phl[(size_t)gj * nx + gi];
        }
        s_phi[k] = p;
        if (HCZ) s_phl[k] = pl;
    }
    __syncthreads();

    // ---- stage B: macros and bar-plus on the one-cell halo ----
    for (int k = tid; k < R1N; k += NTHREADS) {
        const int rx = k % R1X, ry = k / R1X;
        const int gi = wrapc(i0 - 1 + rx, nx);
        int gj = j0 - 1 + ry;
        if (WALLS && (gj < 0 || gj >= ny)) continue;  // wall ghost: extrapolated below
        gj = wrapc(gj, ny);
        const size_t n = (size_t)gj * nx + gi;
        T ht[Q9], ft[Q9];
#pragma unroll
        for (int q = 0; q < Q9; ++q) {
            ht[q] = in.h[(size_t)q * N + n];
            ft[q] = in.f[(size_t)q * N + n];
        }
        const int c2 = (ry + 1) * R2X + (rx + 1);  // this cell in the R2 frame
        T pc[4] = {s_phi[c2 + 1], s_phi[c2 + R2X], s_phi[c2 - 1], s_phi[c2 - R2X]};
        T pd[4] = {s_phi[c2 + R2X + 1], s_phi[c2 + R2X - 1], s_phi[c2 - R2X - 1], s_phi[c2 - R2X + 1]};
        T lc[4] = {0, 0, 0, 0}, ld[4] = {0, 0, 0, 0}, pl = 0, ppu = 0, ppv = 0;
        if (HCZ) {
            lc[0] = s_phl[c2 + 1]; lc[1] = s_phl[c2 + R2X]; lc[2] = s_phl[c2 - 1]; lc[3] = s_phl[c2 - R2X];
            ld[0] = s_phl[c2 + R2X + 1]; ld[1] = s_phl[c2 + R2X - 1];
            ld[2] = s_phl[c2 - R2X - 1]; ld[3] = s_phl[c2 - R2X + 1];
            pl = s_phl[c2];
            ppu = in.ppu[n];
            ppv = in.ppv[n];
        }
        CellMacro<T> m;
        AcSource<T> s;
        T mu;
        derive_cell<T, MODEL, UNIT>(C, ht, ft, pl, lc, ld, ppu, ppv, pc, pd, m, s, mu);
        T hb[Q9], fb[Q9];
        cell_barplus<T, MODEL, UNIT>(C, m, ht, ft, hb, fb);
#pragma unroll
        for (int q = 0; q < Q9; ++q) {
            s_bp[q][k] = hb[q];
            s_bp[Q9 + q][k] = fb[q];
        }
        s_mac[0][k] = m.px; s_mac[1][k] = m.py; s_mac[2][k] = m.rx;
        s_mac[3][k] = m.ry; s_mac[4][k] = m.Fx; s_mac[5][k] = m.Fy;
        const bool interior = rx >= 1 && rx <= TX && ry >= 1 && ry <= TY && i0 - 1 + rx < nx && j0 - 1 + ry < ny;
        if (WALLS == 2 && interior && (gj == 0 || gj == ny - 1)) {
            // Wall_3_NEE values of this cell's wall face (the phase path's
            // expressions: same macros, same tilde populations)
            T *v = s_nee + ((gj == 0 ? 0 : 1) * TX + (rx - 1)) * 2 * Q9;
            const T xc = ((T)gi + (T)0.5) * (T)C.dx;
            nee_face<T, MODEL, UNIT>(C, m, ht, ft, xc, (T)(gj == 0 ? 0 : ny) * (T)C.dx, xc,
                               ((T)gj + (T)0.5) * (T)C.dx, v, v + Q9);
        }
        if (interior && (HCZ || WALLS)) out.phl[n] = m.phi;
        if (interior && HCZ) {
            out.ppu[n] = s.ppu;
            out.ppv[n] = s.ppv;
            if (WALLS && (gj == 0 || gj == ny - 1)) {
                // ghost that goes with the next step's phi_lag stencil
                const int side = gj == 0 ? 0 : 1;
                const int nxt = gj == 0 ? c2 + R2X : c2 - R2X;
                out.gl[side * nx + gi] = (T)2.0 * s_phl[c2] - s_phl[nxt];
            }
        }
    }
    __syncthreads();

    // ---- wall ghost rows of the one-cell halo (ExtrapolationfBP, Ark/Cpp/Boundary.cpp:38-63) ----
    if (WALLS) {
        for (int k = tid; k < R1N; k += NTHREADS) {
            const int rx = k % R1X, ry = k / R1X;
            const int gj = j0 - 1 + ry;
            if (gj != -1 && gj != ny) continue;
            const int dirn = gj < 0 ? 1 : -1;  // towards the fluid
            const int kn = (ry + dirn) * R1X + rx, knn = (ry + 2 * dirn) * R1X + rx;
            const int kg = ry * R1X + rx;
#pragma unroll
            for (int q = 0; q < 2 * Q9; ++q) s_bp[q][kg] = 2 * s_bp[q][kn] - s_bp[q][knn];
#pragma unroll
            for (int a = 0; a < 6; ++a) s_mac[a][kg] = (T)2.0 * s_mac[a][kn] - s_mac[a][knn];
        }
        __syncthreads();
    }

    // ---- stage C: face fluxes ----
    // Each thread keeps the fluxes of its x- and y-face (the six populations
    // with a non-zero normal velocity; the other three carry a zero flux) in
    // registers until every face has read the bar-plus planes, then writes
    // them over those planes.
    // 16 x 16: the faces past one per thread first, on waves 2 (x) and 3 (y),
    // straight into LDS of their own (the face planes overlay the bar-plus
    // planes, which the other faces still read)
    constexpr int EXF = G::EXF, EYF = G::EYF;
    constexpr int MXF = NXF < NTHREADS ? NXF : NTHREADS, MYF = NYF < NTHREADS ? NYF : NTHREADS;
    __shared__ T s_ext[EXF + EYF > 0 ? 12 * (EXF + EYF) : 1];
    if (EXF > 0 && tid >= 128 && tid < 128 + EXF) {
        const int e = tid - 128;
        T h[6], f[6];
        xface64<T, MODEL, UNIT, G>(C, s_bp, s_mac, NTHREADS + e, h, f);
#pragma unroll
        for (int z = 0; z < 6; ++z) { s_ext[z * EXF + e] = h[z]; s_ext[(6 + z) * EXF + e] = f[z]; }
    }
    if (EYF > 0 && tid >= 192 && tid < 192 + EYF) {
        const int e = tid - 192;
        T h[6], f[6];
        yface64<T, MODEL, WALLS, UNIT, G>(C, s_bp, s_mac, s_nee, NTHREADS + e, j0, ny, h, f);
#pragma unroll
        for (int z = 0; z < 6; ++z) {
            s_ext[12 * EXF + z * EYF + e] = h[z];
            s_ext[12 * EXF + (6 + z) * EYF + e] = f[z];
        }
    }
    if (EXF + EYF > 0) __builtin_amdgcn_sched_barrier(0);
    // keep the two face passes apart: interleaved, they hold both faces'
    // operands at once (past 256 VGPRs, one wave per SIMD)
    T xh[6], xf[6], yh[6], yf[6];
    if (tid < MXF) xface64<T, MODEL, UNIT, G>(C, s_bp, s_mac, tid, xh, xf);
    __builtin_amdgcn_sched_barrier(0);
    if (tid < MYF) yface64<T, MODEL, WALLS, UNIT, G>(C, s_bp, s_mac, s_nee, tid, j0, ny, yh, yf);
    // own bar-plus before the region is reused for the face fluxes
    const int ci = tid % TX, cj = tid / TX;
    const bool cell = tid < TX * TY;
    const int kc = (cj + 1) * R1X + ci + 1;
    T hbp[Q9], fbp[Q9];
    if (cell) {
#pragma unroll
        for (int q = 0; q < Q9; ++q) { hbp[q] = s_bp[q][kc]; fbp[q] = s_bp[Q9 + q][kc]; }
    }
    __syncthreads();
    if (tid < MXF) {
#pragma unroll
        for (int z = 0; z < 6; ++z) {
            s_fx[z * NXF + tid] = xh[z];
            s_fx[(6 + z) * NXF + tid] = xf[z];
        }
    }
    if (tid < MYF) {
#pragma unroll
        for (int z = 0; z < 6; ++z) {
            s_fy[z * NYF + tid] = yh[z];
            s_fy[(6 + z) * NYF + tid] = yf[z];
        }
    }
    if (EXF > 0 && tid < EXF) {
#pragma unroll
        for (int z = 0; z < 12; ++z) s_fx[z * NXF + NTHREADS + tid] = s_ext[z * EXF + tid];
    }
    if (EYF > 0 && tid >= 64 && tid < 64 + EYF) {
        const int e = tid - 64;
#pragma unroll
        for (int z = 0; z < 12; ++z) s_fy[z * NYF + NTHREADS + e] = s_ext[12 * EXF + z * EYF + e];
    }
    __syncthreads();

    // ---- stage D: Update_phi_T (Ark/Cpp/DUGKS2DSolver.cpp:748-792), faces E, N, W, S ----
    // (a population with a zero normal velocity on a face contributes a zero
    // flux there; skipping it leaves every sum unchanged)
    const int gi = i0 + ci, gj = j0 + cj;
    if (!cell || gi >= nx || gj >= ny) return;
    const size_t n = (size_t)gj * nx + gi;
    const int kW = cj * (TX + 1) + ci, kE = kW + 1;
    const int kS = cj * TX + ci, kN = kS + TX;
    const T aTP = (T)(4.0 / 3.0), bTP = (T)(1.0 / 3.0), dtv = (T)C.dtv;
    T phi = 0;
#pragma unroll
    for (int q = 0; q < Q9; ++q) {
        const int zx = q == 1 ? 0 : q == 2 ? 1 : q == 4 ? 2 : q == 5 ? 3 : q == 6 ? 4 : q == 8 ? 5 : -1;
        const int zy = q == 2 ? 0 : q == 3 ? 1 : q == 4 ? 2 : q == 6 ? 3 : q == 7 ? 4 : q == 8 ? 5 : -1;
        T s = 0;
        if (zx >= 0) s += -1 * s_fx[zx * NXF + kE];
        if (zy >= 0) s += -1 * s_fy[zy * NYF + kN];
        if (zx >= 0) s += 1 * s_fx[zx * NXF + kW];
        if (zy >= 0) s += 1 * s_fy[zy * NYF + kS];
        T hn = aTP * hbp[q] - bTP * in.h[(size_t)q * N + n] + dtv * s;
        out.h[(size_t)q * N + n] = hn;
        phi += hn;
        T sf = 0;
        if (zx >= 0) sf += -1 * s_fx[(6 + zx) * NXF + kE];
        if (zy >= 0) sf += -1 * s_fy[(6 + zy) * NYF + kN];
        if (zx >= 0) sf += 1 * s_fx[(6 + zx) * NXF + kW];
        if (zy >= 0) sf += 1 * s_fy[(6 + zy) * NYF + kS];
        out.f[(size_t)q * N + n] = aTP * fbp[q] - bTP * in.f[(size_t)q * N + n] + dtv * sf;
    }
    out.phi[n] = phi;
}

// ---- state conversion kernels ----

// explicit (FP64, reference layout) -> derived planes
template <typename T>
__global__ void __launch_bounds__(256) k_to_derived(Consts C, DevState S, Planes<T> P, int which, double p_ref)
{
    int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t N = S.N;
    if (n >= N) return;
    const bool acp = C.model == MPD_MODEL_AC_P;
    if (which == 0) {  // after the first phase step: lag planes
        if (P.phl) P.phl[n] = (T)S.m[MPD_M_PHI * N + n];
        if (P.ppu) P.ppu[n] = (T)S.m[MPD_M_PPU * N + n];
        if (P.ppv) P.ppv[n] = (T)S.m[MPD_M_PPV * N + n];
    } else {           // after the second: populations and phi
        for (int q = 0; q < Q9; ++q) {
            double base = p_ref == 0.0 ? 0.0 : (acp ? (C.w[q] - (q == 0)) * p_ref / C.rt : C.w[q] * p_ref);
            P.h[q * N + n] = (T)S.h[q * N + n];
            P.f[q * N + n] = (T)(S.f[q * N + n] - base);
        }
        P.phi[n] = (T)S.m[MPD_M_PHI * N + n];
        if (P.phd) P.phd[n] = S.m[MPD_M_PHI * N + n];
    }
}

// walls: ghost Phi of the current explicit state (goes with phi_lag's stencil)
template <typename T>
__global__ void __launch_bounds__(256) k_ghost_phi(Consts C, DevState S, T *gl)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * C.nx) return;
    int side = t / C.nx, i = t % C.nx;
    int jn = side == 0 ? 0 : C.ny - 1, jnn = side == 0 ? 1 : C.ny - 2;
    const double *phi = S.m + MPD_M_PHI * S.N;
    gl[side * C.nx + i] = (T)(2.0 * phi[(size_t)jn * C.nx + i] - phi[(size_t)jnn * C.nx + i]);
}

// derived planes -> explicit reference macros (and tilde planes) in S
template <typename T, int MODEL, int WALLS>
__global__ void __launch_bounds__(256) k_materialize(Consts C, DevState S, Planes<T> P, double p_ref)
{
    int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t N = S.N;
    if (n >= N) return;
    const int nx = C.nx, ny = C.ny;
    const int i = (int)(n % nx), j = (int)(n / nx);
    constexpr bool HCZ = MODEL == MPD_MODEL_AC_HCZ;
    auto PHI = [&](int a, int b) -> T {
        a = wrapc(a, nx);
        if (WALLS && (b < 0 || b >= ny)) {
            int jn = b < 0 ? 0 : ny - 1, jnn = b < 0 ? 1 : ny - 2;
            return (T)2.0 * P.phl[(size_t)jn * nx + a] - P.phl[(size_t)jnn * nx + a];
        }
        return P.phi[(size_t)wrapc(b, ny) * nx + a];
    };
    auto PHL = [&](int a, int b) -> T {
        a = wrapc(a, nx);
        if (WALLS && (b < 0 || b >= ny)) return P.gl[(b < 0 ? 0 : nx) + a];
        return P.phl[(size_t)wrapc(b, ny) * nx + a];
    };
    T ht[Q9], ft[Q9];
    for (int q = 0; q < Q9; ++q) {
        ht[q] = P.h[q * N + n];
        ft[q] = P.f[q * N + n];
    }
    T pc[4] = {PHI(i + 1, j), PHI(i, j + 1), PHI(i - 1, j), PHI(i, j - 1)};
    T pd[4] = {PHI(i + 1, j + 1), PHI(i - 1, j + 1), PHI(i - 1, j - 1), PHI(i + 1, j - 1)};
    T lc[4] = {0, 0, 0, 0}, ld[4] = {0, 0, 0, 0}, pl = 0, ppu = 0, ppv = 0;
    if (HCZ) {
        lc[0] = PHL(i + 1, j); lc[1] = PHL(i, j + 1); lc[2] = PHL(i - 1, j); lc[3] = PHL(i, j - 1);
        ld[0] = PHL(i + 1, j + 1); ld[1] = PHL(i - 1, j + 1); ld[2] = PHL(i - 1, j - 1); ld[3] = PHL(i + 1, j - 1);
        pl = P.phl[n];
        ppu = P.ppu[n];
        ppv = P.ppv[n];
    }
    CellMacro<T> m;
    AcSource<T> s;
    T mu;
    derive_cell<T, MODEL>(C, ht, ft, pl, lc, ld, ppu, ppv, pc, pd, m, s, mu);
    double *M = S.m;
    // FP32-bulk: phi and rho from the conserved FP64 phase field
    M[MPD_M_PHI * N + n] = P.phd ? P.phd[n] : (double)m.phi;
    M[MPD_M_RHO * N + n] = P.phd ? C.aphi * P.phd[n] + C.bphi : (double)m.rho;
    M[MPD_M_U * N + n] = (double)m.U;
    M[MPD_M_V * N + n] = (double)m.V;
    M[MPD_M_P * N + n] = (double)m.p + p_ref;
    M[MPD_M_PHIX * N + n] = (double)s.sx;
    M[MPD_M_PHIY * N + n] = (double)s.sy;
    M[MPD_M_RHOX * N + n] = (double)s.rx;
    M[MPD_M_RHOY * N + n] = (double)s.ry;
    M[MPD_M_FX * N + n] = (double)s.Fx;
    M[MPD_M_FY * N + n] = (double)s.Fy;
    M[MPD_M_LAP * N + n] = (double)s.lap;
    M[MPD_M_PPU * N + n] = (double)s.ppu;
    M[MPD_M_PPV * N + n] = (double)s.ppv;
    M[MPD_M_MU * N + n] = (double)mu;
    M[MPD_M_TAU * N + n] = (double)m.tau;
    const bool acp = MODEL == MPD_MODEL_AC_P;
    for (int q = 0; q < Q9; ++q) {
        double base = p_ref == 0.0 ? 0.0 : (acp ? (C.w[q] - (q == 0)) * p_ref / C.rt : C.w[q] * p_ref);
        S.h[q * N + n] = (double)ht[q];
        S.f[q * N + n] = (double)ft[q] + base;
    }
}

template <typename T>
Planes<T> planes_of(FusedState &F, int which)
{
    Planes<T> P;
    T *b = (T *)F.st[which];
    const int64_t N = F.ncell;
    P.h = b;
    P.f = b + 9 * N;
    P.phi = b + 18 * N;
    P.phl = F.nplanes > 19 ? b + 19 * N : nullptr;
    P.ppu = F.nplanes > 20 ? b + 20 * N : nullptr;
    P.ppv = F.nplanes > 21 ? b + 21 * N : nullptr;
    P.gl = F.gl[which] ? (T *)F.gl[which] : nullptr;
    P.phd = F.phd[which];
    return P;
}

template <typename T, int MODEL, int WALLS, bool UNIT, class G>
void launch_step(FusedState &F, const Consts &C, hipStream_t st)
{
    Planes<T> in = planes_of<T>(F, F.cur), out = planes_of<T>(F, 1 - F.cur);
    const int ntx = (C.nx + G::TX - 1) / G::TX, nty = (C.ny + G::TY - 1) / G::TY;
    const int ntiles = ntx * nty;
    const int nblocks = ((ntiles + 7) / 8) * 8;
    hipLaunchKernelGGL((k_fused_step<T, MODEL, WALLS, UNIT, G>), dim3(nblocks), dim3(NTHREADS), 0, st, C, in, out,
                       (int64_t)F.ncell, ntx, ntiles);
    F.cur = 1 - F.cur;
}

template <typename T, int MODEL, int WALLS>
void launch_step(FusedState &F, const Consts &C, hipStream_t st)
{
    // MaSpan 1 and dx a power of two: constant lattice velocities and face
    // gradients by the exact reciprocal of dx (the same results bit for bit)
    if (F.tile == MPD_TILE_16X16) {
        if (C.unit) launch_step<T, MODEL, WALLS, true, Tile64<16, 16>>(F, C, st);
        else launch_step<T, MODEL, WALLS, false, Tile64<16, 16>>(F, C, st);
    } else {
        if (C.unit) launch_step<T, MODEL, WALLS, true, Tile64<16, 15>>(F, C, st);
        else launch_step<T, MODEL, WALLS, false, Tile64<16, 15>>(F, C, st);
    }
}

template <typename T>
void step_t(FusedState &F, const Consts &C, hipStream_t st);

template <>
void step_t<float>(FusedState &F, const Consts &C, hipStream_t st)
{
    Planes<float> a = planes_of<float>(F, F.cur), b = planes_of<float>(F, 1 - F.cur);
    FusedPlanes32 in{a.h, a.f, a.phi, a.phl, a.ppu, a.ppv, a.gl, a.phd},
        out{b.h, b.f, b.phi, b.phl, b.ppu, b.ppv, b.gl, b.phd};
    fused32_step(in, out, C, F.ncell, F.tile, F.cus, st);
    F.cur = 1 - F.cur;
}

template <>
void step_t<double>(FusedState &F, const Consts &C, hipStream_t st)
{
    using T = double;
    if (C.model == MPD_MODEL_AC_HCZ) {
        if (C.nee) launch_step<T, MPD_MODEL_AC_HCZ, 2>(F, C, st);
        else if (C.walls) launch_step<T, MPD_MODEL_AC_HCZ, 1>(F, C, st);
        else launch_step<T, MPD_MODEL_AC_HCZ, 0>(F, C, st);
    } else {
        if (C.nee) launch_step<T, MPD_MODEL_AC_P, 2>(F, C, st);
        else if (C.walls) launch_step<T, MPD_MODEL_AC_P, 1>(F, C, st);
        else launch_step<T, MPD_MODEL_AC_P, 0>(F, C, st);
    }
}

template <typename T>
void materialize_t(FusedState &F, const DevState &S, const Consts &C, hipStream_t st)
{
    Planes<T> P = planes_of<T>(F, F.cur);
    dim3 g((unsigned)((S.N + 255) / 256));
    if (C.model == MPD_MODEL_AC_HCZ) {
        if (C.walls) hipLaunchKernelGGL((k_materialize<T, MPD_MODEL_AC_HCZ, 1>), g, dim3(256), 0, st, C, S, P, F.p_ref);
        else hipLaunchKernelGGL((k_materialize<T, MPD_MODEL_AC_HCZ, 0>), g, dim3(256), 0, st, C, S, P, F.p_ref);
    } else {
        if (C.walls) hipLaunchKernelGGL((k_materialize<T, MPD_MODEL_AC_P, 1>), g, dim3(256), 0, st, C, S, P, F.p_ref);
        else hipLaunchKernelGGL((k_materialize<T, MPD_MODEL_AC_P, 0>), g, dim3(256), 0, st, C, S, P, F.p_ref);
    }
}

template <typename T>
void convert_t(FusedState &F, const DevState &S, const Consts &C, int which, hipStream_t st)
{
    dim3 g((unsigned)((S.N + 255) / 256));
    Planes<T> P = planes_of<T>(F, F.cur);
    if (which == 0 && C.walls && C.model == MPD_MODEL_AC_HCZ)
        hipLaunchKernelGGL((k_ghost_phi<T>), dim3((2 * C.nx + 255) / 256), dim3(256), 0, st, C, S, P.gl);
    else if (which == 1 || which == 2)
        hipLaunchKernelGGL((k_to_derived<T>), g, dim3(256), 0, st, C, S, P, which - 1, F.p_ref);
}

}  // namespace

FusedPlanes32 fused32_planes(FusedState &F, int which)
{
    Planes<float> a = planes_of<float>(F, which);
    return FusedPlanes32{a.h, a.f, a.phi, a.phl, a.ppu, a.ppv, a.gl, a.phd};
}

int fused64_tile_auto(int nx, int ny, int cus)
{
    // two workgroups per CU (LDS); 16 x 16 where its tiles take fewer rounds of
    // those slots than 16 x 15 tiles (e.g. 256 x 1024: 1,024 tiles, two full
    // rounds, against 1,104), else 16 x 15 (fewer instructions per cell)
    const int64_t slots = 2 * (int64_t)(cus > 0 ? cus : 256);
    const int64_t t15 = (int64_t)((nx + 15) / 16) * ((ny + 14) / 15);
    const int64_t t16 = (int64_t)((nx + 15) / 16) * ((ny + 15) / 16);
    return (t16 + slots - 1) / slots < (t15 + slots - 1) / slots ? MPD_TILE_16X16 : MPD_TILE_16X15;
}

bool fused_supported(const mpd_config &cfg)
{
    // FP32-bulk pseudopotential: the FP32 kernels of pp32.hip (D2Q9 with CD and
    // UW faces, D2V16 with CD faces; periodic)
    if (cfg.model == MPD_MODEL_PP_D2Q9) return cfg.precision == MPD_FP32_BULK && cfg.bc_y == MPD_BC_PERIODIC;
    if (cfg.model == MPD_MODEL_PP_D2V16)
        return cfg.precision == MPD_FP32_BULK && cfg.bc_y == MPD_BC_PERIODIC && cfg.flux == MPD_FLUX_CD;
    const bool ac = (cfg.model == MPD_MODEL_AC_HCZ || cfg.model == MPD_MODEL_AC_P) && cfg.nx >= 4 && cfg.ny >= 4;
    // UW faces: the FP32-bulk tile kernel (the FP64 fused kernel is CD only)
    if (cfg.flux == MPD_FLUX_UW) return ac && cfg.precision == MPD_FP32_BULK;
    return ac && cfg.flux == MPD_FLUX_CD;
}

bool fused_needs_priming(const mpd_config &cfg)
{
    return cfg.model != MPD_MODEL_PP_D2Q9 && cfg.model != MPD_MODEL_PP_D2V16;
}

int fused_nplanes(const mpd_config &cfg)
{
    // h~(9) f~(9) phi, + phi_lag (HCZ, or walls for the ghost rows), + ppu ppv (HCZ)
    return cfg.model == MPD_MODEL_AC_HCZ ? 22 : (cfg.bc_y != MPD_BC_PERIODIC ? 20 : 19);
}

int64_t fused32_state_bytes(const mpd_config &cfg)
{
    // per set: the float planes (fused_alloc) and the FP64 phase field
    return 2 * (int64_t)cfg.nx * cfg.ny * (4 * (int64_t)fused_nplanes(cfg) + 8);
}

int fused_alloc(FusedState &F, const DevState &S, const Consts &C, const mpd_config &cfg)
{
    fused_free(F);
    F.ncell = S.N;
    F.fp32 = cfg.precision == MPD_FP32_BULK;
    F.pp = cfg.model == MPD_MODEL_PP_D2Q9 || cfg.model == MPD_MODEL_PP_D2V16;
    if (F.pp) {
        void *p = nullptr;
        int64_t b = 0;
        if (pp32_alloc(F.pp32, &p, &b, C, S.N, cfg.model == MPD_MODEL_PP_D2V16 ? 16 : 9)) return -1;
        F.mem[F.nmem++] = p;
        F.bytes = b;
        F.ready = true;
        return 0;
    }
    F.nplanes = fused_nplanes(cfg);
    // FP32 keeps f as a deviation from the rest equilibrium at p_ref = 1 (p0 of
    // the reference, Ark/Constant/ZeroReference.h:86)
    F.p_ref = F.fp32 ? 1.0 : 0.0;
    const size_t es = F.fp32 ? sizeof(float) : sizeof(double);
    for (int w = 0; w < 2; ++w) {
        void *p = nullptr;
        if (hipMalloc(&p, es * (size_t)F.nplanes * S.N) != hipSuccess) { fused_free(F); return -1; }
        F.mem[F.nmem++] = p;
        F.st[w] = p;
        F.bytes += (int64_t)(es * (size_t)F.nplanes * S.N);
        F.gl[w] = nullptr;
        F.phd[w] = nullptr;
        if (F.fp32) {
            if (hipMalloc(&p, sizeof(double) * (size_t)S.N) != hipSuccess) { fused_free(F); return -1; }
            F.mem[F.nmem++] = p;
            F.phd[w] = (double *)p;
            F.bytes += (int64_t)(sizeof(double) * (size_t)S.N);
        }
        if (C.walls && cfg.model == MPD_MODEL_AC_HCZ) {
            if (hipMalloc(&p, es * 2 * (size_t)C.nx) != hipSuccess) { fused_free(F); return -1; }
            F.mem[F.nmem++] = p;
            F.gl[w] = p;
            F.bytes += (int64_t)(es * 2 * (size_t)C.nx);
        }
    }
    F.cur = 0;
    F.ready = true;
    return 0;
}

void fused_free(FusedState &F)
{
    for (int k = 0; k < F.nmem; ++k) (void)hipFree(F.mem[k]);
    F.nmem = 0;
    F.bytes = 0;
    F.ready = false;
    F.st[0] = F.st[1] = nullptr;
    F.gl[0] = F.gl[1] = nullptr;
    F.phd[0] = F.phd[1] = nullptr;
}

int64_t fused_bytes(const FusedState &F) { return F.bytes; }

void fused_convert(FusedState &F, const DevState &S, const Consts &C, int which, hipStream_t st)
{
    if (F.pp) {
        pp32_from64(F.pp32, S, st);
        F.pp_stepped = false;
        return;
    }
    if (F.fp32) convert_t<float>(F, S, C, which, st);
    else convert_t<double>(F, S, C, which, st);
}

void fused_step(FusedState &F, const DevState &S, const Consts &C, const mpd_config &cfg, int first, hipStream_t st)
{
    (void)S; (void)cfg; (void)first;
    if (F.pp) {
        pp32_step(F.pp32, C, F.pp_stepped, st);
        F.pp_stepped = true;
    }
    else if (F.fp32) step_t<float>(F, C, st);
    else step_t<double>(F, C, st);
}

int fused_materialize(FusedState &F, const DevState &S, const Consts &C, const mpd_config &cfg, hipStream_t st)
{
    (void)cfg;
    if (F.pp) pp32_to64(F.pp32, S, C, F.pp_stepped, st);
    else if (F.fp32) materialize_t<float>(F, S, C, st);
    else materialize_t<double>(F, S, C, st);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace mpd
