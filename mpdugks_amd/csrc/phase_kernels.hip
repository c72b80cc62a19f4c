// Plain FP64 GPU path: one HIP kernel per phase of the reference time step
// (Ark/Cpp/DUGKS2DSolver.cpp:164-328), structure-of-arrays state, explicit
// cell macros exactly as the reference keeps them.  Built with
// -ffp-contract=off so every phase reproduces the reference's FP64 results.
//
// This is the plain drop-in and the correctness anchor for the fused path.
#include "d2v16.h"
#include "device_state.h"
#include "physics.h"

namespace mpd {

__device__ inline int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// macro field `fld` at (i,j); rows -1 and ny are the wall ghosts
__device__ inline double cellm(const DevState &S, const Consts &C, int fld, int i, int j)
{
    i = wrapi(i, C.nx);
    if (j < 0 || j >= C.ny) {
        if (!C.walls) j = wrapi(j, C.ny);
        else return S.gm[((size_t)fld * 2 + (j < 0 ? 0 : 1)) * C.nx + i];
    }
    return S.m[(size_t)fld * S.N + (size_t)j * C.nx + i];
}

__device__ inline double cellbp(const DevState &S, const Consts &C, const double *bp, const double *gbp,
                                int q, int i, int j)
{
    i = wrapi(i, C.nx);
    if (j < 0 || j >= C.ny) {
        if (!C.walls) j = wrapi(j, C.ny);
        else return gbp[((size_t)q * 2 + (j < 0 ? 0 : 1)) * C.nx + i];
    }
    return bp[(size_t)q * S.N + (size_t)j * C.nx + i];
}

// phase 1: equilibria, sources and bar-plus of every cell
template <int MODEL, bool UNIT>
__global__ void __launch_bounds__(256) k_phase1(DevState S, Consts C)
{
    size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (size_t)S.N) return;
    const size_t N = S.N;
    const double *m = S.m;
    CellMacro<double> cm;
    cm.phi = m[MPD_M_PHI * N + n]; cm.rho = m[MPD_M_RHO * N + n];
    cm.U = m[MPD_M_U * N + n];     cm.V = m[MPD_M_V * N + n];
    cm.p = m[MPD_M_P * N + n];     cm.tau = m[MPD_M_TAU * N + n];
    cm.px = m[MPD_M_PHIX * N + n]; cm.py = m[MPD_M_PHIY * N + n];
    cm.rx = m[MPD_M_RHOX * N + n]; cm.ry = m[MPD_M_RHOY * N + n];
    cm.Fx = m[MPD_M_FX * N + n];   cm.Fy = m[MPD_M_FY * N + n];
    double ht[Q9], ft[Q9], hb[Q9], fb[Q9];
#pragma unroll
    for (int q = 0; q < Q9; ++q) {
        ft[q] = S.f[q * N + n];
        ht[q] = MODEL == MPD_MODEL_PP_D2Q9 ? 0.0 : S.h[q * N + n];
    }
    cell_barplus<double, MODEL, UNIT>(C, cm, ht, ft, hb, fb);
#pragma unroll
    for (int q = 0; q < Q9; ++q) {
        S.fbp[q * N + n] = fb[q];
        if (MODEL != MPD_MODEL_PP_D2Q9) S.hbp[q * N + n] = hb[q];
    }
}

// phase 2: wall ghosts (WallShadowC_fBP / ExtrapolationfBP, Ark/Cpp/Boundary.cpp:38-85)
__global__ void __launch_bounds__(256) k_ghosts(DevState S, Consts C)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * C.nx) return;
    int side = t / C.nx, i = t % C.nx;
    int jn = side == 0 ? 0 : C.ny - 1, jnn = side == 0 ? 1 : C.ny - 2;
    for (int fld = 0; fld < MPD_M_COUNT; ++fld) {
        double nr = S.m[(size_t)fld * S.N + (size_t)jn * C.nx + i];
        double nx2 = S.m[(size_t)fld * S.N + (size_t)jnn * C.nx + i];
        double g;
        if (fld == MPD_M_U || fld == MPD_M_V || fld == MPD_M_MU) g = -nr;
        else if (fld == MPD_M_SCG) g = nx2;
        else g = 2.0 * nr - nx2;
        S.gm[((size_t)fld * 2 + side) * C.nx + i] = g;
    }
    for (int q = 0; q < Q9; ++q) {
        size_t a = (size_t)q * S.N + (size_t)jn * C.nx + i, b = (size_t)q * S.N + (size_t)jnn * C.nx + i;
        S.ghbp[((size_t)q * 2 + side) * C.nx + i] = 2 * S.hbp[a] - S.hbp[b];
        S.gfbp[((size_t)q * 2 + side) * C.nx + i] = 2 * S.fbp[a] - S.fbp[b];
    }
}

// phase 3 (UW only): 6-point gradients of bar-plus (Ark/Lib/GradSchemeBasic.h:95-126).
// With walls a wall-row cell's stencil takes the ghost rows, wrapped in x (its
// diagonal across the periodic line is the corner cell PeriodicShadowC_*, an alias
// of the wall shadow cell at the row's other end: Ark/Cpp/Mesh/MeshCartesian.cpp:61-118,
// Ark/Cpp/Mesh/GhostCells.cpp:20-64).
__global__ void __launch_bounds__(256) k_grad(DevState S, Consts C, const double *bp, const double *gbp,
                                              double *gx, double *gy)
{
    size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (size_t)S.N) return;
    int i = (int)(n % C.nx), j = (int)(n / C.nx);
    auto B = [&](int q, int a, int b) { return cellbp(S, C, bp, gbp, q, a, b); };
    for (int q = 0; q < Q9; ++q) {
        double c[4] = {B(q, i + 1, j), B(q, i, j + 1), B(q, i - 1, j), B(q, i, j - 1)};
        double d[4] = {B(q, i + 1, j + 1), B(q, i - 1, j + 1), B(q, i - 1, j - 1), B(q, i + 1, j - 1)};
        double a, b;
        grad9(C, c, d, a, b);
        gx[(size_t)q * S.N + n] = a;
        gy[(size_t)q * S.N + n] = b;
    }
}

// f-bar at a face (CD: Ark/Cpp/DUGKS2DSolver.cpp:604-687; UW: :560-603,688-706)
template <int FLUX, bool UNIT>
__device__ inline double face_value(const DevState &S, const Consts &C, const double *bp, const double *gbp,
                                    const double *bpx, const double *bpy, int dir, int q, int oi, int oj,
                                    int ni, int nj, double xf, double yf)
{
    double bo = cellbp(S, C, bp, gbp, q, oi, oj), bn = cellbp(S, C, bp, gbp, q, ni, nj);
    if (FLUX == MPD_FLUX_CD) {
        if (dir == 0)
            return cd_face_value<double, UNIT>(C, 0, q, bo, bn, cellbp(S, C, bp, gbp, q, ni, nj + 1),
                                         cellbp(S, C, bp, gbp, q, ni, nj - 1),
                                         cellbp(S, C, bp, gbp, q, oi, oj + 1),
                                         cellbp(S, C, bp, gbp, q, oi, oj - 1));
        return cd_face_value<double, UNIT>(C, 1, q, bo, bn, cellbp(S, C, bp, gbp, q, ni + 1, nj),
                                     cellbp(S, C, bp, gbp, q, ni - 1, nj),
                                     cellbp(S, C, bp, gbp, q, oi + 1, oj),
                                     cellbp(S, C, bp, gbp, q, oi - 1, oj));
    }
    double ex = xi_of<UNIT, double>(C, 0, q), ey = xi_of<UNIT, double>(C, 1, q);
    double xin = dir == 0 ? ex : ey;
    double xo = (oi + 0.5) * C.dx, yo = (oj + 0.5) * C.dx, xn = (ni + 0.5) * C.dx, yn = (nj + 0.5) * C.dx;
    size_t ko = (size_t)q * S.N + (size_t)wrapi(oj, C.ny) * C.nx + wrapi(oi, C.nx);
    size_t kn = (size_t)q * S.N + (size_t)wrapi(nj, C.ny) * C.nx + wrapi(ni, C.nx);
    double dxo = xf - C.hdt * ex - xo, dyo = yf - C.hdt * ey - yo;
    double dxn = xf - C.hdt * ex - xn, dyn = yf - C.hdt * ey - yn;
    // a wall face's other side is the wall's shadow cell, whose gradient the
    // reference never computes (its arrays from `new double[]`,
    // Ark/Cpp/Mesh/Mesh_2D.cpp:8-13, hold 0 in its runs): 0
    const bool go = C.walls && oj < 0, gn = C.walls && nj >= C.ny;
    const double gxo = go ? 0.0 : bpx[ko], gyo = go ? 0.0 : bpy[ko];
    const double gxn = gn ? 0.0 : bpx[kn], gyn = gn ? 0.0 : bpy[kn];
    if (xin > 0) return bo + (dxo * gxo + dyo * gyo);
    if (xin < 0) return bn + (dxn * gxn + dyn * gyn);
    // xi.n = 0: the average adds the reference face's owner first, its mesh
    // cell for a face at index 0 of its normal (the zone-12 copy of a periodic
    // face, the bottom wall; Ark/Lib/MeshConstructFunc.h:216-233)
    if (oi < 0 || oj < 0) return (bn + (dxn * gxn + dyn * gyn) + bo + (dxo * gxo + dyo * gyo)) / 2;
    return (bo + (dxo * gxo + dyo * gyo) + bn + (dxn * gxn + dyn * gyn)) / 2;
}

__device__ inline void face_geom(const Consts &C, int dir, int i, int j, int &oi, int &oj, int &ni, int &nj,
                                 double &xf, double &yf)
{
    if (dir == 0) { oi = i - 1; oj = j; ni = i; nj = j; xf = i * C.dx; yf = (j + 0.5) * C.dx; }
    else { oi = i; oj = j - 1; ni = i; nj = j; xf = (i + 0.5) * C.dx; yf = j * C.dx; }
}

// Faces of the second copies: x-faces of column nx (one per row), then, with
// periodic y, y-faces of row ny (one per column).
__host__ __device__ inline int uw_copies(const Consts &C) { return C.ny + (C.walls ? 0 : C.nx); }

// phase 4 + 6 for the Allen-Cahn models: Flux_2D on every face, Wall_3_BB on
// wall faces, times xi.n dS (Ark/Cpp/DUGKS2DSolver.cpp:219-273).  UW: the
// reference stores each periodic face twice (zone 12 owned by the first
// column / row, zone 8 by the last) and the two round apart in the xi.n = 0
// average, so threads past the structured faces evaluate the zone-8 copies
// into S.cbhl / S.cbfl [Q][nx + ny], which the last column / row take.
template <int MODEL, int FLUX, bool UNIT>
__global__ void __launch_bounds__(256) k_faces_ac(DevState S, Consts C)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t nxf = S.N, nyf = (size_t)C.nyf * C.nx;
    const size_t ncopy = FLUX == MPD_FLUX_UW ? (size_t)uw_copies(C) : 0;
    if (t >= nxf + nyf + ncopy) return;
    int dir, i, j;
    size_t fidx;
    const bool copy = t >= nxf + nyf;
    if (copy) {
        fidx = t - nxf - nyf;
        dir = fidx < (size_t)C.ny ? 0 : 1;
        i = dir == 0 ? C.nx : (int)(fidx - C.ny);
        j = dir == 0 ? (int)fidx : C.ny;
    } else {
        dir = t < nxf ? 0 : 1;
        fidx = dir == 0 ? t : t - nxf;
        i = (int)(fidx % C.nx);
        j = (int)(fidx / C.nx);
    }
    int oi, oj, ni, nj;
    double xf, yf;
    face_geom(C, dir, i, j, oi, oj, ni, nj, xf, yf);
    double hb[Q9], fb[Q9];
#pragma unroll
    for (int q = 0; q < Q9; ++q) {
        hb[q] = face_value<FLUX, UNIT>(S, C, S.hbp, S.ghbp, S.hbpx, S.hbpy, dir, q, oi, oj, ni, nj, xf, yf);
        fb[q] = face_value<FLUX, UNIT>(S, C, S.fbp, S.gfbp, S.fbpx, S.fbpy, dir, q, oi, oj, ni, nj, xf, yf);
    }
    FaceMacro<double> fm;
    fm.px = 0.5 * (cellm(S, C, MPD_M_PHIX, oi, oj) + cellm(S, C, MPD_M_PHIX, ni, nj));
    fm.py = 0.5 * (cellm(S, C, MPD_M_PHIY, oi, oj) + cellm(S, C, MPD_M_PHIY, ni, nj));
    fm.rx = 0.5 * (cellm(S, C, MPD_M_RHOX, oi, oj) + cellm(S, C, MPD_M_RHOX, ni, nj));
    fm.ry = 0.5 * (cellm(S, C, MPD_M_RHOY, oi, oj) + cellm(S, C, MPD_M_RHOY, ni, nj));
    fm.Fx = 0.5 * (cellm(S, C, MPD_M_FX, oi, oj) + cellm(S, C, MPD_M_FX, ni, nj));
    fm.Fy = 0.5 * (cellm(S, C, MPD_M_FY, oi, oj) + cellm(S, C, MPD_M_FY, ni, nj));
    face_macros_ac<double, MODEL, UNIT>(C, hb, fb, fm);
    double ah = 2.0 * fm.tau / (2.0 * fm.tau + C.hdt);
    double bh = 1.0 - ah;
    double ch = fm.tau * bh;
    double hh[Q9], ff[Q9];
#pragma unroll
    for (int q = 0; q < Q9; ++q) face_bgk<double, MODEL, UNIT>(C, fm, q, hb[q], fb[q], ah, bh, ch, hh[q], ff[q]);
    if (C.walls && dir == 1 && (j == 0 || j == C.ny) && C.nee) {
        // Wall_3_NEE (Ark/Cpp/Boundary.cpp:140-184) from the fluid cell's
        // macros and tilde populations; bounce-back follows
        const int jc = j == 0 ? 0 : C.ny - 1;
        const size_t N = S.N, n = (size_t)jc * C.nx + i;
        const double *m = S.m;
        CellMacro<double> cm;
        cm.phi = m[MPD_M_PHI * N + n]; cm.rho = m[MPD_M_RHO * N + n];
        cm.U = m[MPD_M_U * N + n];     cm.V = m[MPD_M_V * N + n];
        cm.p = m[MPD_M_P * N + n];     cm.tau = m[MPD_M_TAU * N + n];
        cm.px = m[MPD_M_PHIX * N + n]; cm.py = m[MPD_M_PHIY * N + n];
        cm.rx = m[MPD_M_RHOX * N + n]; cm.ry = m[MPD_M_RHOY * N + n];
        cm.Fx = m[MPD_M_FX * N + n];   cm.Fy = m[MPD_M_FY * N + n];
        double ht[Q9], ft[Q9];
#pragma unroll
        for (int q = 0; q < Q9; ++q) { ht[q] = S.h[q * N + n]; ft[q] = S.f[q * N + n]; }
        nee_face<double, MODEL, UNIT>(C, cm, ht, ft, xf, yf, (i + 0.5) * C.dx, (jc + 0.5) * C.dx, hh, ff);
    }
    if (C.walls && dir == 1 && (j == 0 || j == C.ny)) {
        // Wall_3_BB (Ark/Cpp/Boundary.cpp:118-139)
        double h0[Q9], f0[Q9];
#pragma unroll
        for (int q = 0; q < Q9; ++q) { h0[q] = hh[q]; f0[q] = ff[q]; }
#pragma unroll
        for (int q = 0; q < Q9; ++q) {
            bool in = j == 0 ? (ey9(q) > 0) : (ey9(q) < 0);
            if (in) { hh[q] = h0[bb9(q)]; ff[q] = f0[bb9(q)]; }
        }
    }
    size_t plane = dir == 0 ? nxf : nyf;
    double *hfl = S.hfl + (dir ? (size_t)Q9 * nxf : 0), *ffl = S.ffl + (dir ? (size_t)Q9 * nxf : 0);
    if (copy) {
        plane = (size_t)C.nx + C.ny;
        hfl = S.cbhl;
        ffl = S.cbfl;
    }
#pragma unroll
    for (int q = 0; q < Q9; ++q) {
        double xin = C.dx * (dir == 0 ? C.xi[0][q] : C.xi[1][q]);
        hfl[(size_t)q * plane + fidx] = hh[q] * xin;
        ffl[(size_t)q * plane + fidx] = ff[q] * xin;
    }
}

// With UW faces in the reference face order the second stored copy of each
// periodic face (zone 8, owned by the last column / row) has its own
// reconstruction (the xi.n = 0 average adds its owner first), so its own macros,
// Psi, force and flux: the face passes evaluate those nx + ny objects too.
template <int FLUX>
__device__ inline bool pp_copies(const DevState &S, const Consts &C)
{
    return FLUX == MPD_FLUX_UW && C.face_ref && S.cbpsi != nullptr;
}

// Pseudopotential faces, pass 1: face macros and Psi of every face
// (Update_MacroVar_h, Ark/Lib/D2Q9PP.cpp:110-136; Update_PseudoPsi(face),
// Ark/Cpp/PseudoP.cpp:23-30).  All Psi are written before any face force reads them.
template <int FLUX, bool UNIT>
__global__ void __launch_bounds__(256) k_pp_face_psi(DevState S, Consts C)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t F2 = 2 * (size_t)S.N, nb = (size_t)C.nx + C.ny;
    // threads past the structured faces: the second copies (pp_copies)
    const bool copy = t >= F2;
    if (t >= F2 + (pp_copies<FLUX>(S, C) ? nb : 0)) return;
    int dir, i, j;
    if (copy) {
        const int b = (int)(t - F2);
        dir = b < C.ny ? 0 : 1;
        i = dir == 0 ? C.nx : b - C.ny;
        j = dir == 0 ? b : C.ny;
    } else {
        dir = t < (size_t)S.N ? 0 : 1;
        const size_t fidx = dir == 0 ? t : t - S.N;
        i = (int)(fidx % C.nx);
        j = (int)(fidx / C.nx);
    }
    int oi, oj, ni, nj;
    double xf, yf;
    face_geom(C, dir, i, j, oi, oj, ni, nj, xf, yf);
    double s = 0, sx = 0, sy = 0, fb[Q9];
#pragma unroll
    for (int q = 0; q < Q9; ++q)
        fb[q] = face_value<FLUX, UNIT>(S, C, S.fbp, S.gfbp, S.fbpx, S.fbpy, dir, q, oi, oj, ni, nj, xf, yf);
#pragma unroll
    for (int q = 0; q < Q9; ++q) s += fb[q];
#pragma unroll
    for (int q = 0; q < Q9; ++q) sx += fb[q] * xi_of<UNIT, double>(C, 0, q);
#pragma unroll
    for (int q = 0; q < Q9; ++q) sy += fb[q] * xi_of<UNIT, double>(C, 1, q);
    double rho = s;
    double U = sx / rho, V = sy / rho;
    const size_t b = copy ? t - F2 : 0;  // index of a second copy
    U += 0.5 * C.hdt * (copy ? S.cbffx[b] : S.ffx[t]) / rho;
    V += 0.5 * C.hdt * (copy ? S.cbffy[b] : S.ffy[t]) / rho;
    double p = C.pp_k * carnahan_starling(rho, C.t_cs);
    double scg, psi;
    calc_psi(p, rho, C.rt, scg, psi);
    if (copy) {
        S.cbpsi[b] = psi;
        S.cbscr[0 * nb + b] = rho;
        S.cbscr[1 * nb + b] = U;
        S.cbscr[2 * nb + b] = V;
        S.cbscr[3 * nb + b] = scg;
        return;
    }
    S.fpsi[t] = psi;
    S.fscr[0 * F2 + t] = rho;
    S.fscr[1 * F2 + t] = U;
    S.fscr[2 * F2 + t] = V;
    S.fscr[3 * F2 + t] = scg;
}

// Face objects of the reference's one-thread face order (Consts::face_ref).
// The reference stores each periodic face twice (zones 12 and 8) and visits
// its faces in FaceArray order: interior x-faces row by row, interior y-faces
// row by row, then the left, right, bottom and top periodic zones
// (Ark/Cpp/DUGKS2DSolver.cpp:219-230, for meshes written by
// mpdugks_amd.mesh.write_cas).  An x-face object is (i, j), i in [0, nx]: i = 0
// and i = nx are the two copies of the periodic face; y-faces likewise in j.
// Both copies see the same new/old pattern of neighbours, so their forces stay
// equal and the copy at i = 0 (j = 0) stands for the pair.
__device__ inline int64_t face_rank(const Consts &C, int dir, int i, int j)
{
    const int64_t nx = C.nx, ny = C.ny, NX = ny * (nx - 1), NI = NX + (ny - 1) * nx;
    if (dir == 0) return i == 0 ? NI + j : (i == nx ? NI + ny + j : (int64_t)j * (nx - 1) + (i - 1));
    return j == 0 ? NI + 2 * ny + i : (j == ny ? NI + 2 * ny + nx + i : NX + (int64_t)(j - 1) * nx + i);
}

// Psi of the stencil neighbour at (di, dj) as the object (dir, i, j) of rank r0
// reads it: faceFaces (Ark/Cpp/Mesh/MeshCartesian.cpp:286-368) reach the copy
// on the referencing face's side of a periodic line; a neighbour visited
// earlier gives its new Psi, one visited later the Psi of the previous step.
__device__ inline double psi_seen(const DevState &S, const Consts &C, int dir, int i, int j, int di, int dj,
                                  int64_t r0, bool copies)
{
    int a = i + di, b = j + dj;
    if (dir == 0) {
        a = a < 0 ? a + C.nx : (a > C.nx ? a - C.nx : a);
        b = wrapi(b, C.ny);
    } else {
        a = wrapi(a, C.nx);
        b = b < 0 ? b + C.ny : (b > C.ny ? b - C.ny : b);
    }
    if (copies && ((dir == 0 && a == C.nx) || (dir == 1 && b == C.ny))) {  // a second copy
        const size_t k = dir == 0 ? (size_t)b : (size_t)C.ny + a;
        return face_rank(C, dir, a, b) < r0 ? S.cbpsi[k] : S.cbpsi_old[k];
    }
    const size_t fi = (size_t)dir * S.N + (size_t)(b == C.ny ? 0 : b) * C.nx + (a == C.nx ? 0 : a);
    return face_rank(C, dir, a, b) < r0 ? S.fpsi[fi] : S.fpsi_old[fi];
}

// Pseudopotential faces, pass 2: face force from the face-lattice gradient of
// Psi (update_MQ_x/y(Face_2D*), Ark/Lib/GradSchemeBasic.h:181-208;
// Update_PseudoForce(face), Ark/Cpp/PseudoP.cpp:39-46), then the BGK half step.
template <int FLUX, bool UNIT>
__global__ void __launch_bounds__(256) k_pp_face_flux(DevState S, Consts C)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t F2 = 2 * (size_t)S.N, nb = (size_t)C.nx + C.ny;
    const bool copies = pp_copies<FLUX>(S, C), copy = t >= F2;
    if (t >= F2 + (copies ? nb : 0)) return;
    int dir, i, j;
    size_t fidx = 0;
    if (copy) {
        const int b = (int)(t - F2);
        dir = b < C.ny ? 0 : 1;
        i = dir == 0 ? C.nx : b - C.ny;
        j = dir == 0 ? b : C.ny;
    } else {
        dir = t < (size_t)S.N ? 0 : 1;
        fidx = dir == 0 ? t : t - S.N;
        i = (int)(fidx % C.nx);
        j = (int)(fidx / C.nx);
    }
    double c[4], d[4];
    if (C.face_ref) {
        const int64_t r0 = face_rank(C, dir, i, j);
        c[0] = psi_seen(S, C, dir, i, j, 1, 0, r0, copies);
        c[1] = psi_seen(S, C, dir, i, j, 0, 1, r0, copies);
        c[2] = psi_seen(S, C, dir, i, j, -1, 0, r0, copies);
        c[3] = psi_seen(S, C, dir, i, j, 0, -1, r0, copies);
        d[0] = psi_seen(S, C, dir, i, j, 1, 1, r0, copies);
        d[1] = psi_seen(S, C, dir, i, j, -1, 1, r0, copies);
        d[2] = psi_seen(S, C, dir, i, j, -1, -1, r0, copies);
        d[3] = psi_seen(S, C, dir, i, j, 1, -1, r0, copies);
    } else {
        const double *ps = S.fpsi + (size_t)dir * S.N;
        auto P = [&](int a, int b) { return ps[(size_t)wrapi(b, C.ny) * C.nx + wrapi(a, C.nx)]; };
        c[0] = P(i + 1, j); c[1] = P(i, j + 1); c[2] = P(i - 1, j); c[3] = P(i, j - 1);
        d[0] = P(i + 1, j + 1); d[1] = P(i - 1, j + 1); d[2] = P(i - 1, j - 1); d[3] = P(i + 1, j - 1);
    }
    double gx, gy;
    grad9(C, c, d, gx, gy);
    const size_t b = copy ? t - F2 : 0;  // index of a second copy
    const double *sc = copy ? S.cbscr + b : S.fscr + t;
    const size_t ss = copy ? nb : F2;
    FaceMacro<double> fm;
    fm.rho = sc[0 * ss];
    fm.U = sc[1 * ss];
    fm.V = sc[2 * ss];
    double scg = sc[3 * ss], psi = copy ? S.cbpsi[b] : S.fpsi[t];
    gx *= -scg * psi;
    gy *= -scg * psi;
    fm.Fx = gx;
    fm.Fy = gy;
    if (copy) {
        S.cbffx[b] = gx;
        S.cbffy[b] = gy;
    } else {
        S.ffx[t] = gx;
        S.ffy[t] = gy;
    }
    int oi, oj, ni, nj;
    double xf, yf;
    face_geom(C, dir, i, j, oi, oj, ni, nj, xf, yf);
    double *ffl = copy ? S.cbfl + b : S.ffl + (dir ? (size_t)Q9 * S.N : 0) + fidx;
    const size_t fs = copy ? nb : (size_t)S.N;
#pragma unroll
    for (int q = 0; q < Q9; ++q) {
        double fb = face_value<FLUX, UNIT>(S, C, S.fbp, S.gfbp, S.fbpx, S.fbpy, dir, q, oi, oj, ni, nj, xf, yf);
        double hh, ff;
        face_bgk<double, MPD_MODEL_PP_D2Q9, UNIT>(C, fm, q, 0.0, fb, C.ah_pp, C.bh_pp, C.ch_pp, hh, ff);
        double xin = C.dx * (dir == 0 ? C.xi[0][q] : C.xi[1][q]);
        ffl[(size_t)q * fs] = ff * xin;
    }
}

// phase 8 + 9: Update_phi_T (Ark/Cpp/DUGKS2DSolver.cpp:748-792, faces E,N,W,S)
// and Update_MacroVar of the cell.
template <int MODEL, bool UNIT>
__global__ void __launch_bounds__(256) k_update(DevState S, Consts C)
{
    size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (size_t)S.N) return;
    const size_t N = S.N;
    int i = (int)(n % C.nx), j = (int)(n / C.nx);
    size_t yplane = (size_t)C.nyf * C.nx;
    size_t kE = (size_t)j * C.nx + wrapi(i + 1, C.nx), kW = n;
    int jn = C.walls ? j + 1 : wrapi(j + 1, C.ny);
    size_t kN = (size_t)jn * C.nx + i, kS = n;
    const double aTP = 4.0 / 3.0, bTP = 1.0 / 3.0;
    const double *hx = S.hfl, *hy = S.hfl + (size_t)Q9 * N;
    const double *fx = S.ffl, *fy = S.ffl + (size_t)Q9 * N;
    // UW (Allen-Cahn; pseudopotential in the reference face order): the last
    // column / row owns the second copy of its periodic east / north face
    // (Face_C, Ark/Cpp/Mesh/MeshCartesian.cpp:28-60)
    const bool copies = C.flux == MPD_FLUX_UW && (MODEL != MPD_MODEL_PP_D2Q9 || (C.face_ref && S.cbpsi != nullptr));
    const bool cE = copies && i == C.nx - 1, cN = copies && !C.walls && j == C.ny - 1;
    const size_t nb = (size_t)C.nx + C.ny;
    double ht[Q9], ft[Q9];
#pragma unroll
    for (int q = 0; q < Q9; ++q) {
        if (MODEL != MPD_MODEL_PP_D2Q9) {
            double s = 0.0;
            s += -1 * (cE ? S.cbhl[q * nb + j] : hx[q * N + kE]);
            s += -1 * (cN ? S.cbhl[q * nb + C.ny + i] : hy[q * yplane + kN]);
            s += 1 * hx[q * N + kW];
            s += 1 * hy[q * yplane + kS];
            ht[q] = aTP * S.hbp[q * N + n] - bTP * S.h[q * N + n] + C.dtv * s;
            S.h[q * N + n] = ht[q];
        }
        double s = 0.0;
        s += -1 * (cE ? S.cbfl[q * nb + j] : fx[q * N + kE]);
        s += -1 * (cN ? S.cbfl[q * nb + C.ny + i] : fy[q * yplane + kN]);
        s += 1 * fx[q * N + kW];
        s += 1 * fy[q * yplane + kS];
        ft[q] = aTP * S.fbp[q * N + n] - bTP * S.f[q * N + n] + C.dtv * s;
        S.f[q * N + n] = ft[q];
    }
    double *m = S.m;
    if (MODEL == MPD_MODEL_PP_D2Q9) {
        // Update_MacroVar, Ark/Lib/D2Q9PP.cpp:89-109 + Update_PseudoPsi(cell)
        double s = 0, sx = 0, sy = 0;
#pragma unroll
        for (int q = 0; q < Q9; ++q) s += ft[q];
#pragma unroll
        for (int q = 0; q < Q9; ++q) sx += ft[q] * xi_of<UNIT, double>(C, 0, q);
#pragma unroll
        for (int q = 0; q < Q9; ++q) sy += ft[q] * xi_of<UNIT, double>(C, 1, q);
        double rho = s, U = sx / rho, V = sy / rho;
        U += C.hdt * m[MPD_M_FX * N + n] / rho;
        V += C.hdt * m[MPD_M_FY * N + n] / rho;
        double p = C.pp_k * carnahan_starling(rho, C.t_cs), scg, psi;
        calc_psi(p, rho, C.rt, scg, psi);
        m[MPD_M_RHO * N + n] = rho;
        m[MPD_M_U * N + n] = U;
        m[MPD_M_V * N + n] = V;
        m[MPD_M_P * N + n] = p;
        m[MPD_M_SCG * N + n] = scg;
        m[MPD_M_PSI * N + n] = psi;
        return;
    }
    double phi, rho, U, V, p, mu, tau;
    cell_macrovar_ac<double, MODEL, UNIT>(C, ht, ft, m[MPD_M_FX * N + n], m[MPD_M_FY * N + n],
                                    m[MPD_M_RHOX * N + n], m[MPD_M_RHOY * N + n], phi, rho, U, V, p, mu, tau);
    m[MPD_M_PHI * N + n] = phi;
    m[MPD_M_RHO * N + n] = rho;
    m[MPD_M_U * N + n] = U;
    m[MPD_M_V * N + n] = V;
    m[MPD_M_P * N + n] = p;
    m[MPD_M_MU * N + n] = mu;
    m[MPD_M_TAU * N + n] = tau;
}

// phase 10: MacroSource (Ark/Cpp/SourceM.cpp:35-62)
__global__ void __launch_bounds__(256) k_source(DevState S, Consts C)
{
    size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (size_t)S.N) return;
    const size_t N = S.N;
    int i = (int)(n % C.nx), j = (int)(n / C.nx);
    double c[4] = {cellm(S, C, MPD_M_PHI, i + 1, j), cellm(S, C, MPD_M_PHI, i, j + 1),
                   cellm(S, C, MPD_M_PHI, i - 1, j), cellm(S, C, MPD_M_PHI, i, j - 1)};
    double d[4] = {cellm(S, C, MPD_M_PHI, i + 1, j + 1), cellm(S, C, MPD_M_PHI, i - 1, j + 1),
                   cellm(S, C, MPD_M_PHI, i - 1, j - 1), cellm(S, C, MPD_M_PHI, i + 1, j - 1)};
    double *m = S.m;
    AcSource<double> s;
    ac_source<double>(C, m[MPD_M_PHI * N + n], m[MPD_M_RHO * N + n], m[MPD_M_U * N + n], m[MPD_M_V * N + n],
                      m[MPD_M_PPU * N + n], m[MPD_M_PPV * N + n], c, d, s);
    m[MPD_M_LAP * N + n] = s.lap;
    m[MPD_M_RHOX * N + n] = s.rx;
    m[MPD_M_RHOY * N + n] = s.ry;
    m[MPD_M_FX * N + n] = s.Fx;
    m[MPD_M_FY * N + n] = s.Fy;
    m[MPD_M_PHIX * N + n] = s.sx;
    m[MPD_M_PHIY * N + n] = s.sy;
    m[MPD_M_PPU * N + n] = s.ppu;
    m[MPD_M_PPV * N + n] = s.ppv;
}

// phase 11: Update_PseudoForce(cell) (Ark/Cpp/PseudoP.cpp:31-38); reads Psi only
__global__ void __launch_bounds__(256) k_ppforce(DevState S, Consts C)
{
    size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (size_t)S.N) return;
    const size_t N = S.N;
    int i = (int)(n % C.nx), j = (int)(n / C.nx);
    double c[4] = {cellm(S, C, MPD_M_PSI, i + 1, j), cellm(S, C, MPD_M_PSI, i, j + 1),
                   cellm(S, C, MPD_M_PSI, i - 1, j), cellm(S, C, MPD_M_PSI, i, j - 1)};
    double d[4] = {cellm(S, C, MPD_M_PSI, i + 1, j + 1), cellm(S, C, MPD_M_PSI, i - 1, j + 1),
                   cellm(S, C, MPD_M_PSI, i - 1, j - 1), cellm(S, C, MPD_M_PSI, i + 1, j - 1)};
    double gx, gy;
    grad9(C, c, d, gx, gy);
    double sp = -S.m[MPD_M_SCG * N + n] * S.m[MPD_M_PSI * N + n];
    gx *= sp;
    gy *= sp;
    S.m[MPD_M_FX * N + n] = gx;
    S.m[MPD_M_FY * N + n] = gy;
}

// ---------------- D2V16 pseudopotential (FP64, CD faces, periodic) ----------------
// The D2V16 set (Ark/Lib/D2V16.cpp:11-75) with the pseudopotential force of the
// D2Q9PP model: the phases of the D2Q9PP kernels above with sixteen populations
// and the moment-matrix equilibrium; isothermal at the reference temperature
// (p = rho R0 T0 in Macro16).

constexpr d2v16::Table kV16 = d2v16::make_table();

// f_eq_j = sum_k iArkC[j][k] M16_k (Update_phi_Eq, Ark/Lib/D2V16.cpp:157-166)
__device__ inline void feq_v16(const Consts &C, double rho, double U, double V, double *feq)
{
    double M[d2v16::Q];
    d2v16::macro16<double>(M, rho, U, V, C.rt, rho * C.rt);
#pragma unroll
    for (int j = 0; j < d2v16::Q; ++j) {
        double s = 0;
#pragma unroll
        for (int k = 0; k < d2v16::Q; ++k) s += kV16.a[j][k] * M[k];
        feq[j] = s;
    }
}

// phase 1: equilibria, He-Shan-Doolen source, bar-plus (DUGKS2DSolver.cpp:379-401)
__global__ void __launch_bounds__(256) k_v16_phase1(DevState S, Consts C)
{
    size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (size_t)S.N) return;
    const size_t N = S.N;
    const double *m = S.m;
    const double rho = m[MPD_M_RHO * N + n], U = m[MPD_M_U * N + n], V = m[MPD_M_V * N + n];
    const double Fx = m[MPD_M_FX * N + n], Fy = m[MPD_M_FY * N + n], tau = m[MPD_M_TAU * N + n];
    const double abp = (2.0 * tau - C.hdt) / (2.0 * tau + C.dt);
    const double bbp = 1.0 - abp;
    const double cbp = tau * bbp;
    double feq[d2v16::Q];
    feq_v16(C, rho, U, V, feq);
#pragma unroll
    for (int q = 0; q < d2v16::Q; ++q) {
        const double ex = kV16.x[q], ey = kV16.y[q];
        double fso = (Fx * (ex - U) + Fy * (ey - V)) * feq[q] / (rho * C.rt);
        double b = abp * S.f[q * N + n] + bbp * feq[q];
        b += cbp * fso;
        S.fbp[q * N + n] = b;
    }
}

// CD face value (CD_Interior_phi_Bh, DUGKS2DSolver.cpp:604-687) of population q
__device__ inline double face_value_v16(const DevState &S, const Consts &C, int dir, int q, int oi, int oj, int ni,
                                        int nj)
{
    const double *bp = S.fbp + (size_t)q * S.N;
    auto B = [&](int a, int b) { return bp[(size_t)wrapi(b, C.ny) * C.nx + wrapi(a, C.nx)]; };
    const double bo = B(oi, oj), bn = B(ni, nj);
    const double dn = (bn - bo) / C.dx;
    const double ex = kV16.x[q], ey = kV16.y[q];
    if (dir == 0) {
        double dt = (0.5 * (B(ni, nj + 1) - B(ni, nj - 1)) + 0.5 * (B(oi, oj + 1) - B(oi, oj - 1))) / C.twodx;
        return 0.5 * (bo + bn) - C.hdt * (dn * ex + dt * ey);
    }
    double dt = (0.5 * (B(ni + 1, nj) - B(ni - 1, nj)) + 0.5 * (B(oi + 1, oj) - B(oi - 1, oj))) / C.twodx;
    return 0.5 * (bo + bn) - C.hdt * (dt * ex + dn * ey);
}

// faces, pass 1: face macros and Psi of every face (as k_pp_face_psi)
__global__ void __launch_bounds__(256) k_v16_face_psi(DevState S, Consts C)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * (size_t)S.N) return;
    int dir = t < (size_t)S.N ? 0 : 1;
    size_t fidx = dir == 0 ? t : t - S.N;
    int i = (int)(fidx % C.nx), j = (int)(fidx / C.nx);
    int oi, oj, ni, nj;
    double xf, yf;
    face_geom(C, dir, i, j, oi, oj, ni, nj, xf, yf);
    double fb[d2v16::Q], s = 0, sx = 0, sy = 0;
#pragma unroll
    for (int q = 0; q < d2v16::Q; ++q) fb[q] = face_value_v16(S, C, dir, q, oi, oj, ni, nj);
#pragma unroll
    for (int q = 0; q < d2v16::Q; ++q) s += fb[q];
#pragma unroll
    for (int q = 0; q < d2v16::Q; ++q) sx += fb[q] * kV16.x[q];
#pragma unroll
    for (int q = 0; q < d2v16::Q; ++q) sy += fb[q] * kV16.y[q];
    double rho = s;
    double U = sx / rho, V = sy / rho;
    U += 0.5 * C.hdt * S.ffx[t] / rho;
    V += 0.5 * C.hdt * S.ffy[t] / rho;
    double p = C.pp_k * carnahan_starling(rho, C.t_cs);
    double scg, psi;
    calc_psi(p, rho, C.rt, scg, psi);
    const size_t F2 = 2 * (size_t)S.N;
    S.fpsi[t] = psi;
    S.fscr[0 * F2 + t] = rho;
    S.fscr[1 * F2 + t] = U;
    S.fscr[2 * F2 + t] = V;
    S.fscr[3 * F2 + t] = scg;
}

// faces, pass 2: face force from the face-lattice gradient of Psi, BGK half
// step (Update_phi_h, DUGKS2DSolver.cpp:402-423) and flux (:424-441)
__global__ void __launch_bounds__(256) k_v16_face_flux(DevState S, Consts C)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * (size_t)S.N) return;
    int dir = t < (size_t)S.N ? 0 : 1;
    size_t fidx = dir == 0 ? t : t - S.N;
    int i = (int)(fidx % C.nx), j = (int)(fidx / C.nx);
    const double *ps = S.fpsi + (size_t)dir * S.N;
    auto P = [&](int a, int b) { return ps[(size_t)wrapi(b, C.ny) * C.nx + wrapi(a, C.nx)]; };
    double c[4] = {P(i + 1, j), P(i, j + 1), P(i - 1, j), P(i, j - 1)};
    double d[4] = {P(i + 1, j + 1), P(i - 1, j + 1), P(i - 1, j - 1), P(i + 1, j - 1)};
    double gx, gy;
    grad9(C, c, d, gx, gy);
    const size_t F2 = 2 * (size_t)S.N;
    const double rho = S.fscr[0 * F2 + t], U = S.fscr[1 * F2 + t], V = S.fscr[2 * F2 + t];
    const double scg = S.fscr[3 * F2 + t], psi = S.fpsi[t];
    gx *= -scg * psi;
    gy *= -scg * psi;
    S.ffx[t] = gx;
    S.ffy[t] = gy;
    int oi, oj, ni, nj;
    double xf, yf;
    face_geom(C, dir, i, j, oi, oj, ni, nj, xf, yf);
    double feq[d2v16::Q];
    feq_v16(C, rho, U, V, feq);
    double *ffl = S.ffl + (dir ? (size_t)d2v16::Q * S.N : 0);
#pragma unroll
    for (int q = 0; q < d2v16::Q; ++q) {
        const double ex = kV16.x[q], ey = kV16.y[q];
        double fb = face_value_v16(S, C, dir, q, oi, oj, ni, nj);
        double fso = (gx * (ex - U) + gy * (ey - V)) * feq[q] / (rho * C.rt);
        double r = C.ah_pp * fb + C.bh_pp * feq[q];
        r += C.ch_pp * fso;
        double xin = C.dx * (dir == 0 ? ex : ey);
        ffl[(size_t)q * S.N + fidx] = r * xin;
    }
}

// Update_phi_T (faces E, N, W, S) and the pseudopotential cell macros
// (Ark/Lib/D2Q9PP.cpp:89-109 with sixteen populations; Update_PseudoPsi)
__global__ void __launch_bounds__(256) k_v16_update(DevState S, Consts C)
{
    size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (size_t)S.N) return;
    const size_t N = S.N;
    int i = (int)(n % C.nx), j = (int)(n / C.nx);
    size_t kE = (size_t)j * C.nx + wrapi(i + 1, C.nx), kW = n;
    size_t kN = (size_t)wrapi(j + 1, C.ny) * C.nx + i, kS = n;
    const double aTP = 4.0 / 3.0, bTP = 1.0 / 3.0;
    const double *fx = S.ffl, *fy = S.ffl + (size_t)d2v16::Q * N;
    double ft[d2v16::Q];
#pragma unroll
    for (int q = 0; q < d2v16::Q; ++q) {
        double s = 0.0;
        s += -1 * fx[q * N + kE];
        s += -1 * fy[q * N + kN];
        s += 1 * fx[q * N + kW];
        s += 1 * fy[q * N + kS];
        ft[q] = aTP * S.fbp[q * N + n] - bTP * S.f[q * N + n] + C.dtv * s;
        S.f[q * N + n] = ft[q];
    }
    double s = 0, sx = 0, sy = 0;
#pragma unroll
    for (int q = 0; q < d2v16::Q; ++q) s += ft[q];
#pragma unroll
    for (int q = 0; q < d2v16::Q; ++q) sx += ft[q] * kV16.x[q];
#pragma unroll
    for (int q = 0; q < d2v16::Q; ++q) sy += ft[q] * kV16.y[q];
    double *m = S.m;
    double rho = s, U = sx / rho, V = sy / rho;
    U += C.hdt * m[MPD_M_FX * N + n] / rho;
    V += C.hdt * m[MPD_M_FY * N + n] / rho;
    double p = C.pp_k * carnahan_starling(rho, C.t_cs), scg, psi;
    calc_psi(p, rho, C.rt, scg, psi);
    m[MPD_M_RHO * N + n] = rho;
    m[MPD_M_U * N + n] = U;
    m[MPD_M_V * N + n] = V;
    m[MPD_M_P * N + n] = p;
    m[MPD_M_SCG * N + n] = scg;
    m[MPD_M_PSI * N + n] = psi;
}

static inline unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

template <int MODEL, int FLUX, bool UNIT>
static void phase_step_t(const DevState &S, const Consts &C, hipStream_t st)
{
    hipLaunchKernelGGL((k_phase1<MODEL, UNIT>), dim3(nblk(S.N)), dim3(256), 0, st, S, C);
    if (C.walls) hipLaunchKernelGGL(k_ghosts, dim3(nblk(2 * C.nx)), dim3(256), 0, st, S, C);
    if (FLUX == MPD_FLUX_UW) {
        if (MODEL != MPD_MODEL_PP_D2Q9)
            hipLaunchKernelGGL(k_grad, dim3(nblk(S.N)), dim3(256), 0, st, S, C, (const double *)S.hbp,
                               (const double *)S.ghbp, S.hbpx, S.hbpy);
        hipLaunchKernelGGL(k_grad, dim3(nblk(S.N)), dim3(256), 0, st, S, C, (const double *)S.fbp,
                           (const double *)S.gfbp, S.fbpx, S.fbpy);
    }
    if (MODEL == MPD_MODEL_PP_D2Q9) {
        const size_t nf = 2 * S.N + (FLUX == MPD_FLUX_UW ? (size_t)C.nx + C.ny : 0);
        hipLaunchKernelGGL((k_pp_face_psi<FLUX, UNIT>), dim3(nblk(nf)), dim3(256), 0, st, S, C);
        hipLaunchKernelGGL((k_pp_face_flux<FLUX, UNIT>), dim3(nblk(nf)), dim3(256), 0, st, S, C);
    } else {
        size_t nf = S.N + (size_t)C.nyf * C.nx + (FLUX == MPD_FLUX_UW ? uw_copies(C) : 0);
        hipLaunchKernelGGL((k_faces_ac<MODEL, FLUX, UNIT>), dim3(nblk(nf)), dim3(256), 0, st, S, C);
    }
    hipLaunchKernelGGL((k_update<MODEL, UNIT>), dim3(nblk(S.N)), dim3(256), 0, st, S, C);
    if (MODEL == MPD_MODEL_PP_D2Q9) hipLaunchKernelGGL(k_ppforce, dim3(nblk(S.N)), dim3(256), 0, st, S, C);
    else hipLaunchKernelGGL(k_source, dim3(nblk(S.N)), dim3(256), 0, st, S, C);
}

static void phase_step_v16(const DevState &S, const Consts &C, hipStream_t st)
{
    hipLaunchKernelGGL(k_v16_phase1, dim3(nblk(S.N)), dim3(256), 0, st, S, C);
    hipLaunchKernelGGL(k_v16_face_psi, dim3(nblk(2 * S.N)), dim3(256), 0, st, S, C);
    hipLaunchKernelGGL(k_v16_face_flux, dim3(nblk(2 * S.N)), dim3(256), 0, st, S, C);
    hipLaunchKernelGGL(k_v16_update, dim3(nblk(S.N)), dim3(256), 0, st, S, C);
    hipLaunchKernelGGL(k_ppforce, dim3(nblk(S.N)), dim3(256), 0, st, S, C);
}

// UNIT (Consts::unit): the lattice velocities as constants and 1/dx exact;
// the same results bit for bit as the generic instantiation
template <bool UNIT>
static void phase_step_m(const DevState &S, const Consts &C, hipStream_t st)
{
    switch (C.model * 2 + C.flux) {
    case MPD_MODEL_AC_P * 2 + MPD_FLUX_CD: phase_step_t<MPD_MODEL_AC_P, MPD_FLUX_CD, UNIT>(S, C, st); break;
    case MPD_MODEL_AC_P * 2 + MPD_FLUX_UW: phase_step_t<MPD_MODEL_AC_P, MPD_FLUX_UW, UNIT>(S, C, st); break;
    case MPD_MODEL_AC_HCZ * 2 + MPD_FLUX_CD: phase_step_t<MPD_MODEL_AC_HCZ, MPD_FLUX_CD, UNIT>(S, C, st); break;
    case MPD_MODEL_AC_HCZ * 2 + MPD_FLUX_UW: phase_step_t<MPD_MODEL_AC_HCZ, MPD_FLUX_UW, UNIT>(S, C, st); break;
    case MPD_MODEL_PP_D2Q9 * 2 + MPD_FLUX_CD: phase_step_t<MPD_MODEL_PP_D2Q9, MPD_FLUX_CD, UNIT>(S, C, st); break;
    case MPD_MODEL_PP_D2Q9 * 2 + MPD_FLUX_UW: phase_step_t<MPD_MODEL_PP_D2Q9, MPD_FLUX_UW, UNIT>(S, C, st); break;
    default: break;
    }
}

__global__ void __launch_bounds__(256) k_copies_from_first(DevState S, Consts C, int forces, int psi)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= C.nx + C.ny) return;
    const size_t k = b < C.ny ? (size_t)b * C.nx : (size_t)S.N + (b - C.ny);
    if (forces) {
        S.cbffx[b] = S.ffx[k];
        S.cbffy[b] = S.ffy[k];
    }
    if (psi) S.cbpsi_old[b] = S.fpsi_old[k];
}

void pp_copies_from_first(const DevState &S, const Consts &C, bool forces, bool psi, hipStream_t st)
{
    if (!S.cbpsi_old) return;
    hipLaunchKernelGGL(k_copies_from_first, dim3(nblk((size_t)C.nx + C.ny)), dim3(256), 0, st, S, C, (int)forces,
                       (int)psi);
}

void phase_step(const DevState &S, const Consts &C, hipStream_t st)
{
    if (C.model == MPD_MODEL_PP_D2V16) {
        phase_step_v16(S, C, st);
        return;
    }
    if (C.unit) phase_step_m<true>(S, C, st);
    else phase_step_m<false>(S, C, st);
}

}  // namespace mpd
