// FP32-bulk fused DUGKS step (Allen-Cahn models, CD faces, periodic x; y
// periodic, bounce-back or NEE walls).  Same stages and derived state as the FP64
// fused kernel (fused_kernels.hip), arithmetic rewritten for throughput:
//  - D2Q9 directions and weights are compile-time constants, so e_q . u
//    folds to adds and the populations with zero normal velocity skip the
//    face BGK (their flux is zero);
//  - divisions by per-cell/per-face scalars become one reciprocal each;
//  - the BGK combinations a x + b x_eq + c x_src (a = 1 - b) are written as
//    (1 - b) x + (b x_eq) + c x_src with b folded into the equilibrium terms
//    (one FMA per population on top of the equilibrium), and 1 - b, b the FP32
//    pair whose sum is exactly 1 (split1): the unit coefficient carries no
//    rounding.  Without it the rounding of a + b acts every step on every
//    population, a bias (measured on the Rayleigh-Taylor case: kinetic energy
//    1.2e-3 low after 2 10^4 steps, 6x the spread of FP64 runs perturbed at
//    FP32 round-off);
//  - the D2Q9 weights are rounded so that they sum to exactly 1 in FP32;
//  - AC-p (U == 0 by construction, Ark/Lib/D2Q9AC.cpp:197-198) drops every
//    velocity term and has constant relaxation factors.
// f~ is stored as the deviation from the rest equilibrium at p_ref, so p below
// is p - p_ref (all equilibria and sources are linear in p).
// Compiled with FP contraction enabled.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "fused.h"
#include "physics.h"

namespace mpd {

namespace {

constexpr int NBP = 12;                    // float2 planes: (h, f) bar-plus of q = 0..8, then the
                                           // half macros (ax, ay), (rgx, rgy), (Fx, Fy)
constexpr int NFL = 7;                     // float2 planes per face direction: 6 (h, f) values, (h mass, 0)
constexpr bool PRELOAD_RING = false;       // also hoist the second-pass loads above the phi staging

// Tile of TX x TY interior cells, one interior cell per thread in the update.
//  16x15, 256 threads: 255 x- and 256 y-faces, one of each per thread; the
//    bar-plus ring takes one partial wave after the interior.
//  16x16, 256 threads: 272 faces of each direction (the last 16 of each on one
//    wave in a second pass) and a ring of 68 cells (two waves): about 1.2x the
//    instructions of a 16x15 tile for 1.07x the cells.
//  32x16, 512 threads: 528 x- and 544 y-faces (16 and 32 in the second pass),
//    a ring of 100 cells (two waves): per cell about the work of 16x15, with
//    half as many tiles, in workgroups of twice the LDS (two per CU).
// LDS plane strides in float2 elements: PS (bar-plus) and FS (face values) are
// not multiples of 32 float2 (64 dwords), or the compiler would pair reads of
// two planes at one site into one two-address instruction whose two
// addresses share a bank.
// H is the halo of the bar-plus region: 1 for CD faces; 2 for UW faces, whose
// upwind cell can be a halo cell and needs the 9-point gradient there.  The phi
// region has one cell more (the sources of every bar-plus cell).
template <int TX_, int TY_, int H_ = 1>
struct TileK {
    static constexpr int TX = TX_, TY = TY_, H = H_;
    static constexpr int NT = TX * TY > 256 ? 512 : 256;   // threads
    static constexpr int R1X = TX + 2 * H, R1Y = TY + 2 * H;  // bar-plus region
    static constexpr int R2X = R1X + 2, R2Y = R1Y + 2;        // phi region
    static constexpr int R1N = R1X * R1Y, R2N = R2X * R2Y;
    static constexpr int NIN = TX * TY;                    // interior cells
    static constexpr int NRING = R1N - NIN;                // halo ring of the bar-plus region
    static constexpr int NXF = (TX + 1) * TY;              // x-faces
    static constexpr int NYF = TX * (TY + 1);              // y-faces
    static constexpr int PS = H == 2 ? 388 : (TX == 16 ? 324 : 612);
    static constexpr int FS = TX == 32 ? 548 : (TY == 15 ? 260 : 276);
    static constexpr int EXF = NXF - NT, EYF = NYF - NT;   // second-pass faces
    static constexpr int NE = 32;                          // second-pass slots per direction
    static constexpr int XE0 = NT - 128, YE0 = NT - 64;    // first thread of the x / y second pass
    static constexpr int NBIG = NBP * PS > 2 * NFL * FS ? NBP * PS : 2 * NFL * FS;  // s_big, float2
    static_assert((TX == 16 && (TY == 15 || TY == 16)) || (TX == 32 && TY == 16), "tile shapes");
    static_assert(H == 1 || (H == 2 && TX == 16 && TY == 15), "UW faces: 16x15 tiles");
    static_assert(R1N <= PS && NXF <= FS && NYF <= FS, "plane strides");
    static_assert(PS % 32 != 0 && FS % 32 != 0, "plane strides");
    static_assert(NIN <= NT && NRING - (NT - NIN) <= NT, "stage-B mapping");
    static_assert(EXF <= NE && EYF <= NE && NE <= 64, "second face pass: one wave per direction");
};

struct K32 {
    float ms;        // MaSpan
    float ms_rt;     // MaSpan / RT
    float rt, inv_rt, mlambda0, hdt, hhdt, hhdt_rt, dt, bdt, aphi, bphi, beta, kappa, phi_l, phi_v;
    float tau_a, tau_b;               // tau = tau_a phi + tau_b (HCZ)
    float inv_d2x6, inv_lapden, inv_rtdt, inv_wdphi, kp, dtvf;  // dtvf = dt/V * MaSpan * area
    float ra, rb, rt4, rt2;           // face reconstruction: (1/2 + A, 1/2 - A), tangential factor, 2 rt4
    float gx, gy, rho_ref;
    // BGK factors of the bar-plus (bp) and face half step as  (1 - b) x + b eq + c src,
    // (omb, b) an FP32 pair whose sum is exactly 1 (split1)
    float omb_bp_h, omb_h;            // Allen-Cahn (constant tau_h): 1 - b
    float bpw_h[3], bw_h[3];          // b w_q for the weight classes 0 / axis / diagonal
    float cbpw_h[3], chw_h[3];        // c w_q
    float omb_bp0, omb0;              // AC-p constant factors
    float bpw0[3], bw0[3], cbpw0[3], chw0[3];
    float bprest0;                    // AC-p rest population: b (w_0 - 1) / RT
    float w[3];                       // weights rest / axis / diagonal, w0 = 1 - 4 w1 - 4 w2
    float hneq, fneq0, hdx;           // NEE walls: 2 tau/(2 tau + dt) of h (and of f, AC-p), dx/2
    float uw_n, uw_t, uw_0;           // UW faces: (1/2 - A)/12, A/12, 1/24 (A = hdt MaSpan / dx)
};

// one case of a batched launch (device table entry)
struct BatchCase {
    K32 k;
    int nx, ny;
    int64_t N;
    FusedPlanes32 pl[2];
    int cur;  // plane set holding the current state
};

// periodic wrap; `small` (uniform) selects the general form for meshes narrower
// than the halo span, otherwise one conditional add or subtract suffices
__device__ __forceinline__ int wrap(int i, int n, bool small)
{
    if (small) {
        int r = i % n;
        return r < 0 ? r + n : r;
    }
    i += i < 0 ? n : 0;
    return i >= n ? i - n : i;
}

__device__ __forceinline__ int tile_of_block(int b, int ntiles)
{
    const int nxcd = 8;
    int per = (ntiles + nxcd - 1) / nxcd;
    int t = (b % nxcd) * per + b / nxcd;
    return t < ntiles ? t : -1;
}

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// An LDS index the compiler cannot relate to other indices, so that it does
// not pair the reads of neighbouring sites of one plane into two-address
// instructions (each such pair keeps a base register per plane live).
__device__ __forceinline__ int opaque(int x)
{
    asm("" : "+v"(x));
    return x;
}

// Buffer resources over the planes of one case: loads and stores are a
// per-lane 32-bit byte offset plus a uniform plane offset, with no 64-bit
// address arithmetic (planes of one array span less than 4 GiB).
using rsrc_t = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ rsrc_t make_rsrc(const float *p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float ld(rsrc_t r, uint32_t n, uint32_t plane_bytes = 0)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, n * 4u, plane_bytes, 0));
}
// load at a per-lane byte offset (plane included)
__device__ __forceinline__ float ldb(rsrc_t r, uint32_t bytes)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, bytes, 0u, 0));
}
// Stores take a cache policy SC: 0, or 16 (the SC1 bit, device scope), which
// has the L2 write the state through as it is stored instead of in the
// write-back at the end of the kernel that makes it visible to the next
// step's other XCDs.  That pays where every tile finishes at about the same
// time, a launch of one round of 32x16 tiles (RT 256x1024: 15.85 against
// 16.39 us per step, 512x512: 15.5 against 15.9-16.3, same state bit for bit);
// on launches of several rounds it measured 1-5% slower, and on the 7.8 M-cell
// ensemble batch 18% slower, so they keep 0.  (The streaming NT bit: 17.5 us.)
constexpr int ST_WT = 16;
template <int SC = 0>
__device__ __forceinline__ void st(rsrc_t r, uint32_t n, float v, uint32_t plane_bytes = 0)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, n * 4u, plane_bytes, SC);
}
__device__ __forceinline__ double ld64(rsrc_t r, uint32_t n)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, n * 8u, 0, 0));
}
template <int SC = 0>
__device__ __forceinline__ void st64(rsrc_t r, uint32_t n, double v)
{
    using u2 = unsigned int __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, n * 8u, 0, SC);
}

// The planes of one state set are consecutive in one allocation (h~ 0-8,
// f~ 9-17, phi 18, phi_lag 19, prevPhiU 20, prevPhiV 21; fused_kernels.hip
// planes_of), so one resource covers them and a plane is a uniform offset.
// The lagged ghost-phi rows (walls) and the FP64 phase field are separate arrays.
enum : uint32_t { P_H = 0, P_F = 9, P_PHI = 18, P_PHL = 19, P_PPU = 20, P_PPV = 21 };
struct Rs {
    rsrc_t a, gl, d;
    uint32_t pb;  // bytes per plane
};
__device__ __forceinline__ Rs make_rs(const FusedPlanes32 &p, int64_t N, int nx, int nplanes)
{
    Rs r;
    r.pb = (uint32_t)(N * 4);
    r.a = make_rsrc(p.h, (uint32_t)nplanes * r.pb);
    r.gl = make_rsrc(p.gl ? p.gl : p.h, p.gl ? (uint32_t)(2 * nx * 4) : 0u);
    r.d = make_rsrc((const float *)p.phd, (uint32_t)(N * 8));
    return r;
}

// e_a x + e_b y for compile-time e in {-1, 0, 1} (no multiplications by 0 or 1)
template <int A, int B>
__device__ __forceinline__ float edot(float x, float y)
{
    if (A == 0 && B == 0) return 0.f;
    if (A == 0) return B > 0 ? y : -y;
    if (B == 0) return A > 0 ? x : -x;
    if (A > 0) return B > 0 ? x + y : x - y;
    return B > 0 ? y - x : -(x + y);
}

// LDS layout of the bar-plus region (R1X x R1Y cells).  The LDS serves a
// 64-bit read 16 lanes at a time and a 32-bit read 32 lanes at a time, over 32
// banks; lanes of one group that hit one bank at different addresses serialise.
// 16-wide tiles: interior columns 1..TX at row pitch TX, so that the two 16-lane
// rows of a half-wave fall on disjoint banks, then the halo columns 0 and TX+1
// (two per row).  32-wide tiles: rows of R1X in place, halo columns included, so
// that the 16 lanes of a group read 16 consecutive cells of one row even where
// the first or last of them is a halo cell.  The phi region (R2X x R2Y, 32-bit
// reads of two rows per wave): columns 2..TX+1 at pitch TX, then the four halo
// columns of each row.  Column info of one column: index = row * pitch + off.
struct Col {
    int pitch, off;
};
template <class G>
__device__ __forceinline__ Col col1(int c)
{
    constexpr int TX = G::TX, H = G::H;
    if constexpr (TX == 32) return Col{G::R1X, c};
    const bool in = c >= H && c < H + TX;
    return Col{in ? TX : 2 * H, in ? c - H : TX * G::R1Y + (c < H ? c : c - TX)};
}
template <class G>
__device__ __forceinline__ Col col2(int c)
{
    constexpr int TX = G::TX, H2 = G::H + 1;
    const bool in = c >= H2 && c < H2 + TX;
    return Col{in ? TX : 2 * H2, in ? c - H2 : TX * G::R2Y + (c < H2 ? c : c - TX)};
}
__device__ __forceinline__ int at(Col c, int row) { return row * c.pitch + c.off; }
template <class G>
__device__ __forceinline__ int r1(int col, int row) { return at(col1<G>(col), row); }

// grad(phi), laplacian, SetZero (GradScheme.cpp:347-376) from the phi region s
// at the nine indices of a 3x3 stencil (ix: -1, 0, +1 columns; rows likewise)
struct Stencil {
    int w[3], c[3], e[3];  // rows row-1, row, row+1
};
template <class G>
__device__ __forceinline__ Stencil stencil_idx(int col, int row)
{
    const Col a = col2<G>(col - 1), b = col2<G>(col), d = col2<G>(col + 1);
    Stencil t;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        t.w[r] = at(a, row - 1 + r);
        t.c[r] = at(b, row - 1 + r);
        t.e[r] = at(d, row - 1 + r);
    }
    return t;
}
__device__ __forceinline__ void stencil(const K32 &k, const float *s, const Stencil &t, float &gx, float &gy,
                                        float &lap)
{
    const float e = s[t.e[1]], nn = s[t.c[2]], w = s[t.w[1]], so = s[t.c[0]];
    const float ne = s[t.e[2]], nw = s[t.w[2]], sw = s[t.w[0]], se = s[t.e[0]];
    const float phi = s[t.c[1]];
    gx = (4.f * (e - w) + ((ne - nw) + (se - sw))) * k.inv_d2x6;
    gy = (4.f * (nn - so) + ((ne - se) + (nw - sw))) * k.inv_d2x6;
    lap = (4.f * ((e + nn) + (w + so)) + ((ne + nw) + (sw + se)) - 20.f * phi) * k.inv_lapden;
    if (fabsf(gx) < 1.0E-14f) gx = 0.f;
    if (fabsf(gy) < 1.0E-14f) gy = 0.f;
}

// SourceMomentum (SourceM.cpp:19-34) plus the gravity body force
__device__ __forceinline__ void momentum_source(const K32 &k, float phi, float gx, float gy, float lap, float &Fx,
                                                float &Fy, float &rx, float &ry)
{
    const float mu_c =
        2.f * k.beta * (phi - k.phi_l) * (phi - k.phi_v) * (2.f * phi - k.phi_l - k.phi_v) - k.kappa * lap;
    rx = k.aphi * gx;
    ry = k.aphi * gy;
    const float drho = k.aphi * phi + (k.bphi - k.rho_ref);
    Fx = mu_c * gx - k.gx * drho;
    Fy = mu_c * gy - k.gy * drho;
}

// Quantities of one site (cell centre or face) shared by its populations; the
// weighted equilibrium terms carry the site's BGK factor b, so that pair_eq
// yields b x_eq (the NEE wall values use a copy with factor 1).
struct Site {
    float U, V, G0, UF, UR;            // velocity, -|u|^2 Lambda0, u.F, u.grad rho
    float mFx, mFy, mrx, mry, mAx, mAy;  // MaSpan * (F, grad rho, AC source); AC-p: F / RT
    float wphi[3], wp[3], wr[3];       // b w phi, b w p (AC-p: b w p / RT), b w rho RT per weight class
};

// Equilibria and sources of the opposite pair (q, q') with e_q = (EX, EY),
// e_q' = -e_q (Update_phi_Eq + Update_DVDF_Source: D2Q9HCZ.cpp:54-92,244-257,
// D2Q9AC.cpp:69-105,331-344), the equilibria times the Site's factor b.  fs,
// fsp and hs are the sources without the weight; the caller's factor carries it.
template <int MODEL, int EX, int EY>
__device__ __forceinline__ void pair_eq(const K32 &k, const Site &m, float &heq, float &heqp, float &feq, float &feqp,
                                        float &fs, float &fsp, float &hs)
{
    constexpr int W = (EX == 0 || EY == 0) ? 1 : 2;
    hs = edot<EX, EY>(m.mAx, m.mAy);
    if (MODEL == MPD_MODEL_AC_HCZ) {
        const float u1 = k.ms_rt * edot<EX, EY>(m.U, m.V);
        const float S = fmaf(0.5f * u1, u1, m.G0);
        const float G = S + u1, Gp = S - u1;
        heq = fmaf(m.wphi[W], u1, m.wphi[W]);
        heqp = fmaf(-m.wphi[W], u1, m.wphi[W]);
        feq = fmaf(m.wr[W], G, m.wp[W]);
        feqp = fmaf(m.wr[W], Gp, m.wp[W]);
        const float a = edot<EX, EY>(m.mFx, m.mFy), c = edot<EX, EY>(m.mrx, m.mry);
        const float xF = a - m.UF, xFp = -a - m.UF;
        const float xR = c - m.UR, xRp = -c - m.UR;
        fs = fmaf(G, fmaf(k.rt, xR, xF), xF);      // (1 + G) xF + G RT xR
        fsp = fmaf(Gp, fmaf(k.rt, xRp, xFp), xFp);
    } else {
        heq = heqp = m.wphi[W];
        feq = feqp = m.wp[W];
        fs = edot<EX, EY>(m.mFx, m.mFy);
        fsp = -fs;
    }
}

// Bar-plus of one cell (Update_MacroVar with the lagged sources, MacroSource,
// Update_phi_Eq/Source, Update_phi_BP: D2Q9HCZ.cpp:133-182, D2Q9AC.cpp:146-231,
// SourceM.cpp:35-62, DUGKS2DSolver.cpp:379-401) written to LDS at c.  base
// receives b + (b - x~)/3 of every population (h: q >= 1), the part of
// Update_phi_T (DUGKS2DSolver.cpp:748-792) that does not involve the fluxes.
// Populations (and HCZ prevPhiU/V) of one cell of the bar-plus region, loaded
// ahead of the phi staging so that their latency overlaps it.
struct CellIn {
    float h[9], f[9], pu, pv;
    uint32_t n;
    int gi, gj;
    bool valid;
};
template <int MODEL, int WALLS, int H>
__device__ __forceinline__ void cell_load(const Rs &in, int nx, int ny, bool small, int i0, int j0, int rx, int ry,
                                          CellIn &x)
{
    x.gi = wrap(i0 - H + rx, nx, small);
    int gj = j0 - H + ry;
    x.valid = !(WALLS && (gj < 0 || gj >= ny));  // wall ghost rows: extrapolated afterwards
    x.gj = gj = WALLS ? gj : wrap(gj, ny, small);
    x.n = (uint32_t)(gj * nx + x.gi);
    // a ghost-row cell loads its wall neighbour's populations (unused): every
    // thread issues the same loads, so the waits counted after them stay exact
    const uint32_t nl = x.valid ? x.n : (uint32_t)((gj < 0 ? 0 : ny - 1) * nx + x.gi);
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        x.h[q] = ld(in.a, nl, (P_H + q) * in.pb);
        x.f[q] = ld(in.a, nl, (P_F + q) * in.pb);
    }
    if (MODEL == MPD_MODEL_AC_HCZ) {
        x.pu = ld(in.a, nl, P_PPU * in.pb);
        x.pv = ld(in.a, nl, P_PPV * in.pb);
    }
}

template <int MODEL, int WALLS, class G, int SC>
__device__ __forceinline__ void cell_bp(const K32 &k, const Rs &out, int nx, int ny, int rx, int ry, const CellIn &x,
                                        float2 *S, const float *s_phi, const float *s_phl, float *base, bool interior,
                                        float2 *s_nee)
{
    constexpr bool HCZ = MODEL == MPD_MODEL_AC_HCZ;
    constexpr int TX = G::TX, PS = G::PS;
    if (!x.valid) return;
    const float *h = x.h, *f = x.f;
    const uint32_t n = x.n;
    const int gi = x.gi, gj = x.gj;
    const int c = r1<G>(rx, ry);
    const Stencil sti = stencil_idx<G>(rx + 1, ry + 1);
    const float phi = s_phi[sti.c[1]];
    const float rho = k.aphi * phi + k.bphi;
    Site m;
    float p, bbp, omb, tau = 0.f, rtau = 0.f;
    m.U = m.V = 0.f;
    if (HCZ) {
        float lgx, lgy, llap, lFx, lFy, lrx, lry;
        stencil(k, s_phl, sti, lgx, lgy, llap);
        momentum_source(k, s_phl[sti.c[1]], lgx, lgy, llap, lFx, lFy, lrx, lry);
        const float irr = rcp(rho * k.rt);
        const float sx = (f[1] + f[2] + f[8]) - (f[4] + f[5] + f[6]);
        const float sy = (f[2] + f[3] + f[4]) - (f[6] + f[7] + f[8]);
        m.U = fmaf(k.ms, sx, k.hdt * lFx) * irr;
        m.V = fmaf(k.ms, sy, k.hdt * lFy) * irr;
        const float s = ((f[0] + f[1]) + (f[2] + f[3])) + ((f[4] + f[5]) + (f[6] + f[7])) + f[8];
        p = fmaf(k.hdt * k.rt, lrx * m.U + lry * m.V, s);
        tau = fmaf(phi, k.tau_a, k.tau_b);
        rtau = rcp(2.f * tau + k.dt);
        split1(fmaf(-k.bdt, rtau, 1.f), omb, bbp);   // 1 - b = (2 tau - dt/2)/(2 tau + dt)
    } else {
        p = k.kp * (((f[1] + f[2]) + (f[3] + f[4])) + ((f[5] + f[6]) + (f[7] + f[8])));
        omb = k.omb_bp0;
        bbp = 0.f;
    }
    float gx, gy, lap, Fx, Fy, rgx, rgy;
    stencil(k, s_phi, sti, gx, gy, lap);
    momentum_source(k, phi, gx, gy, lap, Fx, Fy, rgx, rgy);
    float ax, ay, npu = 0.f, npv = 0.f;
    {
        const float L2 = gx * gx + gy * gy;
        float dpu = 0.f, dpv = 0.f;
        if (HCZ) {
            npu = phi * m.U;
            npv = phi * m.V;
            dpu = (npu - x.pu) * k.inv_rtdt;
            dpv = (npv - x.pv) * k.inv_rtdt;
        }
        if (L2 != 0.f) {
            const float t = -4.f * (phi - k.phi_l) * (phi - k.phi_v) * k.inv_wdphi * rsqrtf(L2);
            ax = fmaf(gx, t, dpu);
            ay = fmaf(gy, t, dpv);
        } else {
            ax = dpu;
            ay = dpv;
        }
    }
    m.mAx = k.ms * ax;
    m.mAy = k.ms * ay;
    float cw[3], chw[3];
    if (HCZ) {
        m.G0 = k.mlambda0 * (m.U * m.U + m.V * m.V);
        m.UF = m.U * Fx + m.V * Fy;
        m.UR = m.U * rgx + m.V * rgy;
        m.mFx = k.ms * Fx;
        m.mFy = k.ms * Fy;
        m.mrx = k.ms * rgx;
        m.mry = k.ms * rgy;
        const float r = rho * k.rt;
        const float cbp = tau * bbp, bp = bbp * p, br = bbp * r;
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            m.wphi[t] = k.bpw_h[t] * phi;
            m.wp[t] = k.w[t] * bp;
            m.wr[t] = k.w[t] * br;
            cw[t] = cbp * k.w[t];
            chw[t] = k.cbpw_h[t];
        }
    } else {
        m.G0 = m.UF = m.UR = m.mrx = m.mry = 0.f;
        m.mFx = k.ms * Fx * k.inv_rt;
        m.mFy = k.ms * Fy * k.inv_rt;
        const float pr = p * k.inv_rt;
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            m.wphi[t] = k.bpw_h[t] * phi;
            m.wp[t] = k.bpw0[t] * pr;
            m.wr[t] = 0.f;
            cw[t] = k.cbpw0[t];
            chw[t] = k.cbpw_h[t];
        }
    }
    // NEE walls use the equilibria themselves (factor 1)
    Site mu = m;
    if (WALLS == 2) {
        const float pu = HCZ ? p : p * k.inv_rt, ru = HCZ ? rho * k.rt : 0.f;
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            mu.wphi[t] = k.w[t] * phi;
            mu.wp[t] = k.w[t] * pu;
            mu.wr[t] = k.w[t] * ru;
        }
    }
    // NEE walls (Wall_3_NEE, Boundary.cpp:140-184; WALLS == 2): the macros of a
    // wall face are extrapolated from its fluid cell -- phi along the
    // Allen-Cahn source vector (the reference's Phi_x/y at that point), p of
    // the cell, U = V = 0 -- so the face equilibria are w phi_f and w p (AC-p
    // w p / RT).  A population leaving the fluid takes that plus the cell's
    // non-equilibrium part 2 tau/(2 tau + dt) (x~ - x_eq + dt/2 x_so);
    // bounce-back gives its partner the same value.  Bottom wall (cell row 0):
    // the populations with e_y < 0; top wall (row ny - 1): e_y > 0.
    const bool top = gj == ny - 1;
    const float phf = fmaf(top ? k.hdx : -k.hdx, ay, phi);
    const float pe = HCZ ? p : p * k.inv_rt;
    const float fneq = HCZ ? 2.f * tau * rtau : k.fneq0;
    float nh[3], nf[3];  // pairs (3,7), (2,6), (4,8)
    float bh[9], bf[9];
    {
        // rest population: u1 = 0, G = G0, no Allen-Cahn source
        const float heq = m.wphi[0];
        float feq, fs;
        if (HCZ) {
            feq = fmaf(m.wr[0], m.G0, m.wp[0]);
            fs = -fmaf(m.G0, fmaf(k.rt, m.UR, m.UF), m.UF);
        } else {
            feq = k.bprest0 * p;
            fs = 0.f;
        }
        bh[0] = fmaf(k.omb_bp_h, h[0], heq);
        bf[0] = fmaf(omb, f[0], feq) + cw[0] * fs;
    }
#define MPD_CELL_PAIR(Q, EX, EY)                                                         \
    {                                                                                    \
        constexpr int QP = Q + 4, W = (EX == 0 || EY == 0) ? 1 : 2;                      \
        float heq, heqp, feq, feqp, fs, fsp, hs;                                         \
        pair_eq<MODEL, EX, EY>(k, m, heq, heqp, feq, feqp, fs, fsp, hs);                 \
        bh[Q] = fmaf(k.omb_bp_h, h[Q], heq) + chw[W] * hs;                               \
        bh[QP] = fmaf(k.omb_bp_h, h[QP], heqp) - chw[W] * hs;                            \
        bf[Q] = fmaf(omb, f[Q], feq) + cw[W] * fs;                                       \
        bf[QP] = fmaf(omb, f[QP], feqp) + cw[W] * fsp;                                   \
        if (WALLS == 2 && Q >= 2) {                                                      \
            constexpr int NK = Q == 3 ? 0 : (Q == 2 ? 1 : 2);                            \
            pair_eq<MODEL, EX, EY>(k, mu, heq, heqp, feq, feqp, fs, fsp, hs);            \
            const float hw = k.hdt * k.w[W];                                             \
            const float hx = top ? h[Q] - heq : h[QP] - heqp;                            \
            const float fx = top ? f[Q] - feq : f[QP] - feqp;                            \
            nh[NK] = fmaf(k.w[W], phf, k.hneq * fmaf(hw, top ? hs : -hs, hx));           \
            nf[NK] = fmaf(k.w[W], pe, fneq * fmaf(hw, top ? fs : fsp, fx));              \
        }                                                                                \
    }
    MPD_CELL_PAIR(1, 1, 0)
    MPD_CELL_PAIR(2, 1, 1)
    MPD_CELL_PAIR(3, 0, 1)
    MPD_CELL_PAIR(4, -1, 1)
#undef MPD_CELL_PAIR
    const float third = 1.f / 3.f;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        S[q * PS + c] = make_float2(bh[q], bf[q]);
        if (q > 0) base[q - 1] = fmaf(third, bh[q] - h[q], bh[q]);
        base[8 + q] = fmaf(third, bf[q] - f[q], bf[q]);
    }
    // face averages need (o + n) / 2: store the halves
    S[9 * PS + c] = make_float2(0.5f * ax, 0.5f * ay);
    S[10 * PS + c] = make_float2(0.5f * rgx, 0.5f * rgy);
    S[11 * PS + c] = make_float2(0.5f * Fx, 0.5f * Fy);
    if (WALLS == 2 && interior && (gj == 0 || top)) {
#pragma unroll
        for (int z = 0; z < 3; ++z) s_nee[(top ? 3 * TX : 0) + 3 * (rx - G::H) + z] = make_float2(nh[z], nf[z]);
    }
    if (interior) {
        if (HCZ || WALLS) st<SC>(out.a, n, phi, P_PHL * out.pb);
        if (HCZ) {
            st<SC>(out.a, n, npu, P_PPU * out.pb);
            st<SC>(out.a, n, npv, P_PPV * out.pb);
            if (WALLS && (gj == 0 || gj == ny - 1)) {
                const int nxt = gj == 0 ? sti.c[2] : sti.c[0];
                st<SC>(out.gl, (uint32_t)((gj == 0 ? 0 : nx) + gi), 2.f * s_phl[sti.c[1]] - s_phl[nxt]);
            }
        }
    }
}

// Face reconstruction of one population (CD_Interior_phi_Bh,
// DUGKS2DSolver.cpp:604-687): 1/2 (o + n) - h e_n (n - o)/dx - h e_t (dt o + dt n)/2.
// A population with e_n = 0 carries no flux and enters only the face moments;
// its value is returned doubled (o + n - h e_t (dt o + dt n)), and the moments
// take the 1/2 once.
struct FaceIdx {
    int o, n, op, om, np, nm;  // owner, neighbour, and their tangential neighbours (+/-)
};
template <int DIR, int Q>
__device__ __forceinline__ float2 recon(const K32 &k, const float2 *b, const FaceIdx &x)
{
    constexpr int EN = DIR == 0 ? ex9(Q) : ey9(Q), ET = DIR == 0 ? ey9(Q) : ex9(Q);
    const float2 bo = b[x.o], bn = b[x.n];
    float2 v;
    if (EN > 0) {
        v.x = fmaf(k.rb, bn.x, k.ra * bo.x);
        v.y = fmaf(k.rb, bn.y, k.ra * bo.y);
    } else if (EN < 0) {
        v.x = fmaf(k.ra, bn.x, k.rb * bo.x);
        v.y = fmaf(k.ra, bn.y, k.rb * bo.y);
    } else {
        v.x = bo.x + bn.x;
        v.y = bo.y + bn.y;
    }
    if (ET != 0) {
        const float2 np = b[x.np], nm = b[x.nm], op = b[x.op], om = b[x.om];
        const float tx = (np.x - nm.x) + (op.x - om.x), ty = (np.y - nm.y) + (op.y - om.y);
        const float r = EN == 0 ? k.rt2 : k.rt4;
        const float c = ET > 0 ? -r : r;
        v.x = fmaf(c, tx, v.x);
        v.y = fmaf(c, ty, v.y);
    }
    return v;
}

// Upwind faces (UW_Interior_phi_Bh + Update_phi_fBh, DUGKS2DSolver.cpp:560-603,
// 688-706; gradients Grad_VS_6points, GradSchemeBasic.h:95-126):
// f(x_f) = b_c + (x_f - h xi - x_c) . grad b_c with c the upwind cell (owner for
// e_n > 0, neighbour for e_n < 0, the mean of both for e_n = 0), and
// grad b = (4 (E - W) + (NE - NW + SE - SW), 4 (N - S) + (NE - SE + NW - SW)) / (12 dx).
// The two cells' 3x3 neighbourhoods: x-face 4 columns x 3 rows (a[c * 3 + r],
// owner centre (1, 1), neighbour (2, 1)); y-face 3 columns x 4 rows
// (a[c * 4 + r], owner (1, 1), neighbour (1, 2)).
struct FaceIdxUW {
    int o, n;
    int a[12];
};
template <int DIR, class G>
__device__ __forceinline__ FaceIdxUW face_idx_uw(int oc, int orow)
{
    FaceIdxUW x;
    constexpr int NC = DIR == 0 ? 4 : 3, NR = DIR == 0 ? 3 : 4;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const Col cc = col1<G>(oc - 1 + c);
#pragma unroll
        for (int r = 0; r < NR; ++r) x.a[c * NR + r] = opaque(at(cc, orow - 1 + r));
    }
    x.o = x.a[1 * NR + 1];
    x.n = DIR == 0 ? x.a[2 * NR + 1] : x.a[1 * NR + 2];
    return x;
}
// normal and tangential gradient sums (x 12 dx) of one plane around the
// centre (CC, CR) of a face neighbourhood
template <int DIR, int CC, int CR>
__device__ __forceinline__ void grad_uw(const float2 *b, const FaceIdxUW &x, float2 &gn, float2 &gt)
{
    constexpr int NR = DIR == 0 ? 3 : 4;
    auto v = [&](int dc, int dr) { return b[x.a[(CC + dc) * NR + (CR + dr)]]; };
    const float2 e = v(1, 0), w = v(-1, 0), n = v(0, 1), s = v(0, -1);
    const float2 ne = v(1, 1), nw = v(-1, 1), sw = v(-1, -1), se = v(1, -1);
    float2 gx, gy;
    gx.x = 4.f * (e.x - w.x) + ((ne.x - nw.x) + (se.x - sw.x));
    gx.y = 4.f * (e.y - w.y) + ((ne.y - nw.y) + (se.y - sw.y));
    gy.x = 4.f * (n.x - s.x) + ((ne.x - se.x) + (nw.x - sw.x));
    gy.y = 4.f * (n.y - s.y) + ((ne.y - se.y) + (nw.y - sw.y));
    gn = DIR == 0 ? gx : gy;
    gt = DIR == 0 ? gy : gx;
}
// wall < 0 (> 0): a y-face on the bottom (top) wall, whose owner (neighbour) is
// the wall's shadow cell; its gradient is 0, as in the reference, which never
// computes the gradients of shadow cells (Grad_VS_6points runs over the mesh
// cells, DUGKS2DSolver.cpp:197-206)
template <int DIR, int Q, int WALLS>
__device__ __forceinline__ float2 recon_uw(const K32 &k, const float2 *b, const FaceIdxUW &x, int wall)
{
    constexpr int EN = DIR == 0 ? ex9(Q) : ey9(Q), ET = DIR == 0 ? ey9(Q) : ex9(Q);
    constexpr int OC = 1, OR = 1, NC = DIR == 0 ? 2 : 1, NR = DIR == 0 ? 1 : 2;
    const float ct = ET == 0 ? 0.f : (ET > 0 ? -k.uw_t : k.uw_t);   // -h xi_t / (12 dx)
    const bool zo = WALLS && DIR == 1 && wall < 0, zn = WALLS && DIR == 1 && wall > 0;
    const float2 zero = make_float2(0.f, 0.f);
    float2 v;
    if (EN != 0) {
        float2 gn, gt;
        if (EN > 0) grad_uw<DIR, OC, OR>(b, x, gn, gt);
        else grad_uw<DIR, NC, NR>(b, x, gn, gt);
        if (EN > 0 ? zo : zn) gn = gt = zero;
        const float2 c = b[EN > 0 ? x.o : x.n];
        const float cn = EN > 0 ? k.uw_n : -k.uw_n;                  // (x_f - h xi_n - x_c) / (12 dx)
        v.x = fmaf(cn, gn.x, c.x);
        v.y = fmaf(cn, gn.y, c.y);
        if (ET != 0) {
            v.x = fmaf(ct, gt.x, v.x);
            v.y = fmaf(ct, gt.y, v.y);
        }
    } else {
        float2 gno, gto, gnn, gtn;
        grad_uw<DIR, OC, OR>(b, x, gno, gto);
        grad_uw<DIR, NC, NR>(b, x, gnn, gtn);
        if (zo) gno = gto = zero;
        if (zn) gnn = gtn = zero;
        const float2 bo = b[x.o], bn = b[x.n];
        float2 vo, vn;
        vo.x = fmaf(k.uw_0, gno.x, bo.x);
        vo.y = fmaf(k.uw_0, gno.y, bo.y);
        vn.x = fmaf(-k.uw_0, gnn.x, bn.x);
        vn.y = fmaf(-k.uw_0, gnn.y, bn.y);
        if (ET != 0) {
            vo.x = fmaf(ct, gto.x, vo.x);
            vo.y = fmaf(ct, gto.y, vo.y);
            vn.x = fmaf(ct, gtn.x, vn.x);
            vn.y = fmaf(ct, gtn.y, vn.y);
        }
        v.x = vo.x + vn.x;  // doubled, as in recon
        v.y = vo.y + vn.y;
    }
    return v;
}

// Indices of the face between R1 cells (oc, orow) and its neighbour in
// direction DIR, with the tangential neighbours of both (LDS layout above).
template <int DIR, class G>
__device__ __forceinline__ FaceIdx face_idx(int oc, int orow)
{
    FaceIdx x;
    if (DIR == 0) {
        const Col a = col1<G>(oc), b = col1<G>(oc + 1);
        x.o = at(a, orow);
        x.n = opaque(at(b, orow));
        x.op = opaque(at(a, orow + 1));
        x.om = opaque(at(a, orow - 1));
        x.np = opaque(at(b, orow + 1));
        x.nm = opaque(at(b, orow - 1));
    } else {
        const Col a = col1<G>(oc - 1), b = col1<G>(oc), d = col1<G>(oc + 1);
        x.o = at(b, orow);
        x.n = opaque(at(b, orow + 1));
        x.op = opaque(at(d, orow));
        x.om = opaque(at(a, orow));
        x.np = opaque(at(d, orow + 1));
        x.nm = opaque(at(a, orow + 1));
    }
    return x;
}

// One face: reconstruction, face macros (Update_MacroVar_h, D2Q9HCZ.cpp:183-242,
// D2Q9AC.cpp:232-329), BGK half step (Update_phi_h, DUGKS2DSolver.cpp:402-423)
// and micro-flux (Update_phiFlux_h, :424-441) of the six populations with a
// non-zero normal velocity, in pair order (first member e_n = +1).
template <int DIR, int MODEL, int WALLS, int FLUX, class G, class X>
__device__ __forceinline__ void face(const K32 &k, const float2 *S, const X &x, int wall, float *hfl,
                                     float *ffl, float &mass)
{
    constexpr bool HCZ = MODEL == MPD_MODEL_AC_HCZ;
    constexpr int PS = G::PS;
    const int ko = x.o;
    float hv[9], fv[9];
#define MPD_RECON(Q)                                                                             \
    {                                                                                            \
        float2 v;                                                                                \
        if constexpr (FLUX == MPD_FLUX_UW) v = recon_uw<DIR, Q, WALLS>(k, S + Q * PS, x, wall);  \
        else v = recon<DIR, Q>(k, S + Q * PS, x);                                                \
        hv[Q] = v.x;                                                                             \
        fv[Q] = v.y;                                                                             \
    }
    MPD_RECON(0) MPD_RECON(1) MPD_RECON(2) MPD_RECON(3) MPD_RECON(4)
    MPD_RECON(5) MPD_RECON(6) MPD_RECON(7) MPD_RECON(8)
#undef MPD_RECON
    const int kn = x.n;
    const float2 A0 = S[9 * PS + ko], A1 = S[9 * PS + kn];
    const float2 R0 = S[10 * PS + ko], R1 = S[10 * PS + kn];
    const float2 F0 = S[11 * PS + ko], F1 = S[11 * PS + kn];
    const float ax = A0.x + A1.x, ay = A0.y + A1.y;
    const float rgx = R0.x + R1.x, rgy = R0.y + R1.y;
    const float Fx = F0.x + F1.x, Fy = F0.y + F1.y;
    // hv, fv of the e_n = 0 populations (0 and the pair EA, EB along the face)
    // are doubled; the axis pair along the normal is (NA, NB)
    constexpr int EA = DIR == 0 ? 3 : 1, EB = DIR == 0 ? 7 : 5, NA = DIR == 0 ? 1 : 3, NB = DIR == 0 ? 5 : 7;
    const float phi = fmaf(0.5f, (hv[0] + hv[EA]) + hv[EB], ((hv[NA] + hv[NB]) + (hv[2] + hv[6])) + (hv[4] + hv[8]));
    Site m;
    m.mAx = k.ms * ax;
    m.mAy = k.ms * ay;
    float omb, cw[3];
    if (HCZ) {
        const float rho = k.aphi * phi + k.bphi;
        const float irr = rcp(rho * k.rt);
        // the diagonal populations' part of e_x f and e_y f, then the axis pairs
        const float dx = (fv[2] + fv[8]) - (fv[4] + fv[6]), dy = (fv[2] + fv[4]) - (fv[6] + fv[8]);
        const float sx = DIR == 0 ? dx + (fv[1] - fv[5]) : fmaf(0.5f, fv[1] - fv[5], dx);
        const float sy = DIR == 0 ? fmaf(0.5f, fv[3] - fv[7], dy) : dy + (fv[3] - fv[7]);
        m.U = fmaf(k.ms, sx, k.hhdt * Fx) * irr;
        m.V = fmaf(k.ms, sy, k.hhdt * Fy) * irr;
        const float s = fmaf(0.5f, (fv[0] + fv[EA]) + fv[EB], ((fv[NA] + fv[NB]) + (fv[2] + fv[6])) + (fv[4] + fv[8]));
        const float p = fmaf(k.hhdt_rt, rgx * m.U + rgy * m.V, s);
        const float tau = fmaf(phi, k.tau_a, k.tau_b);
        float bh;
        split1(fmaf(-k.hdt, rcp(2.f * tau + k.hdt), 1.f), omb, bh);   // 1 - b = 2 tau / (2 tau + h)
        const float ch = tau * bh;
        m.G0 = k.mlambda0 * (m.U * m.U + m.V * m.V);
        m.UF = m.U * Fx + m.V * Fy;
        m.UR = m.U * rgx + m.V * rgy;
        m.mFx = k.ms * Fx;
        m.mFy = k.ms * Fy;
        m.mrx = k.ms * rgx;
        m.mry = k.ms * rgy;
        const float r = rho * k.rt;
        const float bp = bh * p, br = bh * r;
#pragma unroll
        for (int t = 1; t < 3; ++t) {
            m.wphi[t] = k.bw_h[t] * phi;
            m.wp[t] = k.w[t] * bp;
            m.wr[t] = k.w[t] * br;
            cw[t] = ch * k.w[t];
        }
    } else {
        const float p = k.kp * fmaf(0.5f, fv[EA] + fv[EB], ((fv[NA] + fv[NB]) + (fv[2] + fv[6])) + (fv[4] + fv[8]));
        omb = k.omb0;
        m.U = m.V = m.G0 = m.UF = m.UR = m.mrx = m.mry = 0.f;
        m.mFx = k.ms * Fx * k.inv_rt;
        m.mFy = k.ms * Fy * k.inv_rt;
        const float pr = p * k.inv_rt;
#pragma unroll
        for (int t = 1; t < 3; ++t) {
            m.wphi[t] = k.bw_h[t] * phi;
            m.wp[t] = k.bw0[t] * pr;
            m.wr[t] = 0.f;
            cw[t] = k.chw0[t];
        }
    }
    // pairs: x-faces (1,5) (2,6) (8,4); y-faces (3,7) (2,6) (4,8)
    float ho[6], fo[6];
#define MPD_FACE_PAIR(Z, Q, QP, EX, EY)                                                  \
    {                                                                                    \
        constexpr int W = (EX == 0 || EY == 0) ? 1 : 2;                                  \
        float heq, heqp, feq, feqp, fs, fsp, hs;                                         \
        pair_eq<MODEL, EX, EY>(k, m, heq, heqp, feq, feqp, fs, fsp, hs);                 \
        ho[Z] = fmaf(k.omb_h, hv[Q], heq) + k.chw_h[W] * hs;                             \
        ho[Z + 1] = fmaf(k.omb_h, hv[QP], heqp) - k.chw_h[W] * hs;                       \
        fo[Z] = fmaf(omb, fv[Q], feq) + cw[W] * fs;                                      \
        fo[Z + 1] = fmaf(omb, fv[QP], feqp) + cw[W] * fsp;                               \
    }
    if (DIR == 0) {
        MPD_FACE_PAIR(0, 1, 5, 1, 0)
        MPD_FACE_PAIR(2, 2, 6, 1, 1)
        MPD_FACE_PAIR(4, 8, 4, 1, -1)
    } else {
        MPD_FACE_PAIR(0, 3, 7, 0, 1)
        MPD_FACE_PAIR(2, 2, 6, 1, 1)
        MPD_FACE_PAIR(4, 4, 8, -1, 1)
    }
#undef MPD_FACE_PAIR
    if (WALLS == 1 && DIR == 1 && wall != 0) {
        // Wall_3_BB (Boundary.cpp:118-139): the populations entering the fluid
        // (e_y > 0 at the bottom wall, < 0 at the top) take their partner's value
#pragma unroll
        for (int z = 0; z < 6; z += 2) {
            if (wall < 0) { ho[z] = ho[z + 1]; fo[z] = fo[z + 1]; }
            else { ho[z + 1] = ho[z]; fo[z + 1] = fo[z]; }
        }
    }
    // the flux is value * MaSpan * area * e_n: the factor MaSpan * area is
    // applied with dt/V in the cell update, the sign e_n there and here
    mass = ((ho[0] - ho[1]) + (ho[2] - ho[3])) + (ho[4] - ho[5]);
#pragma unroll
    for (int z = 0; z < 6; ++z) {
        hfl[z] = ho[z];
        ffl[z] = fo[z];
    }
}

// x-face t of a tile: column t % TX, row t / TX for t < TX*TY, then the last
// column (face index TX*TY + row)
template <int MODEL, int WALLS, int FLUX, class G>
__device__ __forceinline__ void xface(const K32 &k, const float2 *S, int t, float *h, float *f, float &m)
{
    constexpr int NIN = G::NIN, TX = G::TX, H = G::H;
    const int fi = t < NIN ? t % TX : TX, fj = t < NIN ? t / TX : t - NIN;
    // owner: bar-plus region column fi + H - 1, row fj + H
    if constexpr (FLUX == MPD_FLUX_UW)
        face<0, MODEL, WALLS, FLUX, G>(k, S, face_idx_uw<0, G>(fi + H - 1, fj + H), 0, h, f, m);
    else
        face<0, MODEL, WALLS, FLUX, G>(k, S, face_idx<0, G>(fi + H - 1, fj + H), 0, h, f, m);
}

// y-face t of a tile: column t % TX, row t / TX
template <int MODEL, int WALLS, int FLUX, class G>
__device__ __forceinline__ void yface(const K32 &k, const float2 *S, const float2 *s_nee, int t, int j0, int ny,
                                      float *h, float *f, float &m)
{
    constexpr int TX = G::TX, H = G::H;
    const int fi = t % TX, fj = t / TX;
    const int gj = j0 + fj;
    const int wall = WALLS ? (gj == 0 ? -1 : (gj == ny ? 1 : 0)) : 0;
    // owner: bar-plus region column fi + H, row fj + H - 1
    if constexpr (FLUX == MPD_FLUX_UW)
        face<1, MODEL, WALLS, FLUX, G>(k, S, face_idx_uw<1, G>(fi + H, fj + H - 1), wall, h, f, m);
    else
        face<1, MODEL, WALLS, FLUX, G>(k, S, face_idx<1, G>(fi + H, fj + H - 1), wall, h, f, m);
    if (WALLS == 2 && wall != 0) {
        // NEE wall face: the values of the wall-row cell (stage B), both
        // members of each pair alike, so no phase-field mass crosses
        const float2 *v = s_nee + (wall < 0 ? 0 : 3 * TX) + 3 * fi;
#pragma unroll
        for (int z = 0; z < 3; ++z) {
            h[2 * z] = h[2 * z + 1] = v[z].x;
            f[2 * z] = f[2 * z + 1] = v[z].y;
        }
        m = 0.f;
    }
}

// One 16 x TY tile of one case.
//  A: phi (and lagged phi) with a two-cell halo into LDS;
//  B: bar-plus and half macros of the (TX+2) x (TY+2) region (interior cells
//     first, so each interior thread keeps its own b + (b - x~)/3 in registers);
//  C: every x- and y-face of the tile (one of each per thread, then for TY = 16
//     the last 16 of each direction on waves 2 and 3);
//  D: Update_phi_T, phi moved by the antisymmetric face mass fluxes.
// position (rx, ry) of cell r of the halo ring of the bar-plus region
template <class G>
__device__ __forceinline__ void ring_pos(int r, int &rx, int &ry)
{
    constexpr int R1X = G::R1X, R1Y = G::R1Y, TY = G::TY, H = G::H;
    if constexpr (H == 1) {
        // bottom row, top row, left column, right column
        if (r < R1X) { rx = r; ry = 0; }
        else if (r < 2 * R1X) { rx = r - R1X; ry = R1Y - 1; }
        else if (r < 2 * R1X + TY) { rx = 0; ry = r - 2 * R1X + 1; }
        else { rx = R1X - 1; ry = r - 2 * R1X - TY + 1; }
    } else {
        // the H bottom rows, the H top rows, then the 2H halo columns of each other row
        if (r < H * R1X) { ry = r / R1X; rx = r % R1X; }
        else if (r < 2 * H * R1X) { r -= H * R1X; ry = R1Y - H + r / R1X; rx = r % R1X; }
        else {
            r -= 2 * H * R1X;
            const int c = r % (2 * H);
            ry = H + r / (2 * H);
            rx = c < H ? c : R1X - 2 * H + c;
        }
    }
}

template <int MODEL, int WALLS, int FLUX, class G, int SC = 0>
__device__ __forceinline__ void tile32(const K32 &k, int nx, int ny, const FusedPlanes32 &pin,
                                       const FusedPlanes32 &pout, int64_t N, int i0, int j0)
{
    constexpr bool HCZ = MODEL == MPD_MODEL_AC_HCZ;
    constexpr int H = G::H, H2 = G::H + 1;
    static_assert((FLUX == MPD_FLUX_UW) == (H == 2), "halo of the face scheme");
    constexpr int TX = G::TX, TY = G::TY, NT = G::NT, R1X = G::R1X, R1Y = G::R1Y, R2Y = G::R2Y, R2N = G::R2N;
    constexpr int NIN = G::NIN, NRING = G::NRING, NXF = G::NXF, NYF = G::NYF, PS = G::PS, FS = G::FS;
    __shared__ float2 s_big[G::NBIG];
    __shared__ float s_phi[R2N];
    __shared__ float s_phl[HCZ ? R2N : 1];
    __shared__ float2 s_nee[WALLS == 2 ? 6 * TX : 1];  // NEE wall values: [bottom, top][column][pair]
    const int tid = threadIdx.x;
    const bool small = nx < TX + 2 * H2 || ny < TY + 2 * H2;
    constexpr int NPL = HCZ ? 22 : (WALLS ? 20 : 19);
    const Rs in = make_rs(pin, N, nx, NPL), out = make_rs(pout, N, nx, NPL);

    // stage-B cells: interior cells first (thread = cell of stage D), then the
    // ring: rest of the bottom row, top row, left and right columns
    const int ci = tid % TX, cj = tid / TX;
    const bool cell = tid < NIN;
    const bool inside = cell && i0 + ci < nx && j0 + cj < ny;
    constexpr int NR2 = NRING - (NT - NIN);   // ring cells of the second pass
    CellIn x1, x2;
    int rx1, ry1, rx2 = 0, ry2 = 0;
    if (cell) { rx1 = ci + H; ry1 = cj + H; }
    else if constexpr (H == 1) { rx1 = tid - NIN; ry1 = 0; }   // the first NT - NIN < R1X ring cells: bottom row
    else ring_pos<G>(tid - NIN, rx1, ry1);
    // ---- stage A ----
    // phi and phi_lag of up to two cells of the phi region per thread, loaded
    // before this thread's stage-B populations: loads return in order, so the
    // phi staging below waits only for these, and the phi-only part of stage B
    // (stencils, sources) can run while the populations are still arriving
    static_assert(R2N <= 2 * NT, "two phi-region cells per thread");
    // The same loads on every path (no branch around a load), so that the wait
    // before the staging counts only them.  A thread without a second cell
    // loads its first one again.  Wall ghost rows j = -1 and ny take 2 near -
    // next of phi_lag and the stored lagged ghost row; rows beyond take 0.
    float pA[2] = {0.f, 0.f}, plA[2] = {0.f, 0.f};
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int q = tid + u * NT < R2N ? tid + u * NT : tid;
        // LDS index q -> phi-region column and row (layout of col2)
        int rx, ry;
        if (q < TX * R2Y) {
            ry = q / TX;
            rx = q % TX + H2;
        } else {
            const int t = q - TX * R2Y, hc = t % (2 * H2);
            ry = t / (2 * H2);
            rx = hc < H2 ? hc : hc + TX;
        }
        const int gi = wrap(i0 - H2 + rx, nx, small);
        const int gj = j0 - H2 + ry;
        if (WALLS) {
            const bool ghost = gj < 0 || gj >= ny, edge = gj == -1 || gj == ny;
            const int jn = gj < 0 ? 0 : ny - 1, jnn = gj < 0 ? 1 : ny - 2;
            const int j1 = ghost ? jn : gj, j2 = ghost ? jnn : gj;
            const uint32_t o1 = (uint32_t)(j1 * nx + gi) * 4u + (ghost ? P_PHL : P_PHI) * in.pb;
            const uint32_t o2 = (uint32_t)(j2 * nx + gi) * 4u + P_PHL * in.pb;
            const float a1 = ldb(in.a, o1), a2 = ldb(in.a, o2);
            const float a3 = HCZ ? ld(in.gl, (uint32_t)((gj < 0 ? 0 : nx) + gi)) : 0.f;
            pA[u] = ghost ? (edge ? 2.f * a1 - a2 : 0.f) : a1;
            if (HCZ) plA[u] = ghost ? (edge ? a3 : 0.f) : a2;
        } else {
            const uint32_t n = (uint32_t)(wrap(gj, ny, small) * nx + gi);
            pA[u] = ld(in.a, n, P_PHI * in.pb);
            if (HCZ) plA[u] = ld(in.a, n, P_PHL * in.pb);
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    cell_load<MODEL, WALLS, G::H>(in, nx, ny, small, i0, j0, rx1, ry1, x1);
    if (tid < NR2) {
        ring_pos<G>(tid + (NT - NIN), rx2, ry2);
        if (PRELOAD_RING) cell_load<MODEL, WALLS, G::H>(in, nx, ny, small, i0, j0, rx2, ry2, x2);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int q = tid + u * NT;
        if (q >= R2N) break;
        s_phi[q] = pA[u];
        if (HCZ) s_phl[q] = plA[u];
    }
    __syncthreads();

    // ---- stage B ----
    float base[17];
    cell_bp<MODEL, WALLS, G, SC>(k, out, nx, ny, rx1, ry1, x1, s_big, s_phi, s_phl, base, inside, s_nee);
    // 16x16 and 32x16: the waves with a second pass (ring cells here, faces in
    // stage C) are the critical path of their stage; issue them first
    constexpr bool PRIO = G::EXF > 0 || G::EYF > 0;
    if (tid < NR2) {
        if (PRIO) __builtin_amdgcn_s_setprio(2);
        if (!PRELOAD_RING) cell_load<MODEL, WALLS, G::H>(in, nx, ny, small, i0, j0, rx2, ry2, x2);
        float dummy[17];
        cell_bp<MODEL, WALLS, G, SC>(k, out, nx, ny, rx2, ry2, x2, s_big, s_phi, s_phl, dummy, false, s_nee);
        if (PRIO) __builtin_amdgcn_s_setprio(0);
    }
    __syncthreads();

    if (WALLS) {
        // ghost rows (j = -1 below the bottom wall, j = ny above the top wall):
        // WallShadowC_fBP / ExtrapolationfBP (Boundary.cpp:38-117), 2 near - next
        const int gy0 = j0 == 0 ? H - 1 : -1;                   // R1 row of j = -1
        const int gy1 = ny - j0 + H < R1Y ? ny - j0 + H : -1;   // R1 row of j = ny
        const int row = tid < R1X ? gy0 : (tid < 2 * R1X ? gy1 : -1);
        if (row >= 0) {
            const int rx = tid < R1X ? tid : tid - R1X;
            const int dn = row == gy0 && tid < R1X ? 1 : -1;
            const Col cc = col1<G>(rx);
            const int c = at(cc, row), kn = at(cc, row + dn), knn = at(cc, row + 2 * dn);
#pragma unroll
            for (int q = 0; q < NBP; ++q) {
                const float2 a = s_big[q * PS + kn], b = s_big[q * PS + knn];
                s_big[q * PS + c] = make_float2(2.f * a.x - b.x, 2.f * a.y - b.y);
            }
        }
        __syncthreads();
    }

    // ---- stage C ----
    // faces beyond NT (16x16, 32x16) first, on the second-to-last (x) and
    // last (y) wave, straight into LDS of their own (the face planes overlay
    // the bar-plus planes, which the other faces still read); then one x- and
    // one y-face per thread, held in registers until every face is computed
    constexpr int EXF = G::EXF, EYF = G::EYF, NE = G::NE, XE0 = G::XE0, YE0 = G::YE0;
    __shared__ float2 s_ext[EXF > 0 || EYF > 0 ? 2 * NFL * NE : 1];
    int et = -1;
    if (PRIO && tid >= XE0) __builtin_amdgcn_s_setprio(2);  // the last two waves
    if (EXF > 0 && tid >= XE0 && tid < XE0 + EXF) {
        et = tid - XE0;
        float h[6], f[6], m = 0.f;
        xface<MODEL, WALLS, FLUX, G>(k, s_big, NT + et, h, f, m);
#pragma unroll
        for (int z = 0; z < 6; ++z) s_ext[z * NE + et] = make_float2(h[z], f[z]);
        s_ext[6 * NE + et] = make_float2(m, 0.f);
    }
    if (EYF > 0 && tid >= YE0 && tid < YE0 + EYF) {
        et = tid - YE0;
        float h[6], f[6], m = 0.f;
        yface<MODEL, WALLS, FLUX, G>(k, s_big, s_nee, NT + et, j0, ny, h, f, m);
#pragma unroll
        for (int z = 0; z < 6; ++z) s_ext[(NFL + z) * NE + et] = make_float2(h[z], f[z]);
        s_ext[(NFL + 6) * NE + et] = make_float2(m, 0.f);
    }
    float xh[6], xf[6], xm = 0.f, yh[6], yf[6], ym = 0.f;
    if (tid < NXF) xface<MODEL, WALLS, FLUX, G>(k, s_big, tid, xh, xf, xm);
    // NXF >= NT: every thread has both faces, and the scheduler would
    // interleave the two (past 128 VGPRs, below 4 waves per SIMD); keep them apart
    if (NXF >= NT) __builtin_amdgcn_sched_barrier(0);
    if (tid < NYF) yface<MODEL, WALLS, FLUX, G>(k, s_big, s_nee, tid, j0, ny, yh, yf, ym);
    if (PRIO) __builtin_amdgcn_s_setprio(0);
    __syncthreads();
    float2 *s_fx = s_big, *s_fy = s_big + NFL * FS;
    if (tid < NXF) {
#pragma unroll
        for (int z = 0; z < 6; ++z) s_fx[z * FS + tid] = make_float2(xh[z], xf[z]);
        s_fx[6 * FS + tid] = make_float2(xm, 0.f);
    }
    if (tid < NYF) {
#pragma unroll
        for (int z = 0; z < 6; ++z) s_fy[z * FS + tid] = make_float2(yh[z], yf[z]);
        s_fy[6 * FS + tid] = make_float2(ym, 0.f);
    }
    if ((EXF > 0 || EYF > 0) && et >= 0) {
        // this thread's second-pass face into its slot NT + et
        const bool y = tid >= YE0;
        const float2 *src = s_ext + (y ? NFL * NE : 0) + et;
        float2 *d = (y ? s_fy : s_fx) + NT + et;
#pragma unroll
        for (int z = 0; z < NFL; ++z) d[z * FS] = src[z * NE];
    }
    __syncthreads();

    // ---- stage D: x~' = b + (b - x~)/3 + dt/V (W - E + S - N) ----
    if (!inside) return;
    const uint32_t n = (uint32_t)((j0 + cj) * nx + (i0 + ci));
    const int kW = cj * TX + ci, kE = opaque(ci < TX - 1 ? kW + 1 : NIN + cj);
    const int kS = cj * TX + ci, kN = opaque(kS + TX);
    // plane z of the x- (y-) face values holds population xp[z] (yp[z]); even
    // z have e_n = +1, odd z e_n = -1
    constexpr int xp[6] = {1, 5, 2, 6, 8, 4};
    constexpr int yp[6] = {3, 7, 2, 6, 4, 8};
    // every population but 0 has an x- or a y-face term or both: the x terms
    // (populations 1, 5 and the diagonals) and those of 3 and 7 are assigned,
    // the diagonals' y terms added; population 0 has none
    float dh[9], df[9];
#pragma unroll
    for (int z = 0; z < 6; ++z) {
        const int a = (z & 1) ? kE : kW, b = (z & 1) ? kW : kE;
        const float2 va = s_fx[z * FS + a], vb = s_fx[z * FS + b];
        dh[xp[z]] = va.x - vb.x;
        df[xp[z]] = va.y - vb.y;
    }
#pragma unroll
    for (int z = 0; z < 6; ++z) {
        const int a = (z & 1) ? kN : kS, b = (z & 1) ? kS : kN;
        const float2 va = s_fy[z * FS + a], vb = s_fy[z * FS + b];
        if (z < 2) {
            dh[yp[z]] = va.x - vb.x;
            df[yp[z]] = va.y - vb.y;
        } else {
            dh[yp[z]] += va.x - vb.x;
            df[yp[z]] += va.y - vb.y;
        }
    }
    // phi is the conserved variable, kept in FP64: it moves only by the
    // antisymmetric face mass fluxes (each face's sum over populations, shared
    // by both cells, differenced in FP64), and h~_0 takes the remainder so that
    // sum_q h~ = phi.  In exact arithmetic this is the same update; here it
    // keeps the rounding of the per-population products, and of an FP32 phi,
    // from drifting the phase-field mass.  The FP32 copy feeds the stencils.
    const double dm = ((double)s_fx[6 * FS + kW].x - (double)s_fx[6 * FS + kE].x) +
                      ((double)s_fy[6 * FS + kS].x - (double)s_fy[6 * FS + kN].x);
    const double phd = ld64(in.d, n) + (double)k.dtvf * dm;
    const float phi = (float)phd;
    float rest = 0.f;
#pragma unroll
    for (int q = 1; q < 9; ++q) {
        const float hn = fmaf(k.dtvf, dh[q], base[q - 1]);
        st<SC>(out.a, n, hn, (P_H + q) * out.pb);
        rest += hn;
    }
    st<SC>(out.a, n, (float)(phd - (double)rest), P_H * out.pb);
    st64<SC>(out.d, n, phd);
    st<SC>(out.a, n, base[8], P_F * out.pb);
#pragma unroll
    for (int q = 1; q < 9; ++q) st<SC>(out.a, n, fmaf(k.dtvf, df[q], base[8 + q]), (P_F + q) * out.pb);
    st<SC>(out.a, n, phi, P_PHI * out.pb);
}

template <int MODEL, int WALLS, int TX, int TY, int FLUX = MPD_FLUX_CD, int SC = 0>
__global__ void __launch_bounds__(TX * TY > 256 ? 512 : 256) k_fused32(K32 k, int nx, int ny, FusedPlanes32 in,
                                                                        FusedPlanes32 out, int64_t N, int ntx,
                                                                        int ntiles)
{
    using G = TileK<TX, TY, FLUX == MPD_FLUX_UW ? 2 : 1>;
    static_assert((TX * TY > 256 ? 512 : 256) == G::NT, "launch bounds");
    const int tile = tile_of_block(blockIdx.x, ntiles);
    if (tile < 0) return;
    tile32<MODEL, WALLS, FLUX, G, SC>(k, nx, ny, in, out, N, (tile % ntx) * TX, (tile / ntx) * TY);
}

// Batched step: every workgroup takes one (case, tile) entry of a work list
// built on the host; the cases share the model, the y boundary and the tile height.
// Four waves per SIMD (at most 128 VGPRs: two 512-thread or four 256-thread
// workgroups per CU), which the register allocator would otherwise exceed on the
// periodic 32x16 instantiation; NEE batches (16x15 tiles) are left unbounded.
template <int MODEL, int WALLS, int TX, int TY>
__global__ void __launch_bounds__(TX * TY > 256 ? 512 : 256, WALLS == 2 ? 1 : 4)
    k_fused32_batch(const BatchCase *cases, const int2 *work, int nwork, int parity)
{
    const int b = blockIdx.x;
    if (b >= nwork) return;
    const int2 w = work[b];
    if (w.x < 0) return;
    const BatchCase &c = cases[w.x];
    const int sel = c.cur ^ parity;
    const int ntx = (c.nx + TX - 1) / TX;
    tile32<MODEL, WALLS, MPD_FLUX_CD, TileK<TX, TY>>(c.k, c.nx, c.ny, c.pl[sel], c.pl[sel ^ 1], c.N,
                                                     (w.y % ntx) * TX, (w.y / ntx) * TY);
}

K32 make_k32(const Consts &C)
{
    K32 k;
    const double ms = C.xi[0][1];
    const double dnu = C.nu_l - C.nu_v;
    k.ms = (float)ms;
    k.ms_rt = (float)(ms / C.rt);
    k.rt = (float)C.rt;
    k.inv_rt = (float)(1.0 / C.rt);
    k.mlambda0 = (float)(-C.lambda0);
    k.hdt = (float)C.hdt;
    k.hhdt = (float)(0.5 * C.hdt);
    k.hhdt_rt = (float)(0.5 * C.hdt * C.rt);
    k.dt = (float)C.dt;
    k.bdt = (float)(1.5 * C.dt);
    k.aphi = (float)C.aphi;
    k.bphi = (float)C.bphi;
    k.beta = (float)C.beta;
    k.kappa = (float)C.kappa;
    k.phi_l = (float)C.phi_l;
    k.phi_v = (float)C.phi_v;
    // tau = (nu_V + (phi - phi_V)(nu_L - nu_V)) / RT (calcMu/calcTau, D2Q9HCZ.cpp:24-44)
    k.tau_a = (float)(dnu / C.rt);
    k.tau_b = (float)((C.nu_v - C.phi_v * dnu) / C.rt);
    k.inv_d2x6 = (float)(1.0 / C.d2x6);
    k.inv_lapden = (float)(1.0 / C.lapden);
    k.inv_rtdt = (float)(1.0 / C.rtdt);
    k.inv_wdphi = (float)(1.0 / C.wdphi);
    k.kp = (float)C.kp;
    k.dtvf = (float)(C.dtv * ms * C.dx);
    const double A = C.hdt * ms / C.dx;  // normal difference factor h e_n / dx
    k.ra = (float)(0.5 + A);
    k.rb = (float)(0.5 - A);
    k.rt4 = (float)(C.hdt * ms / (2.0 * C.twodx));
    k.rt2 = 2.f * k.rt4;
    k.gx = (float)C.gx;
    k.gy = (float)C.gy;
    k.rho_ref = (float)C.rho_ref;
    // FP32 weights that sum to 1 exactly: w1 = fl(1/9) (a multiple of 2^-27),
    // w2 = 1/36 rounded to a multiple of 2^-27 (one ulp below fl(1/36)), so
    // that w0 = 1 - 4 w1 - 4 w2 is a multiple of 2^-25 in [1/4, 1/2): exact in
    // FP32 (computed in FP64, where every step is exact)
    const double w1d = (double)(1.f / 9.f), w2d = std::ldexp(std::nearbyint(std::ldexp(1.0 / 36.0, 27)), -27);
    k.w[0] = (float)(1.0 - 4.0 * w1d - 4.0 * w2d);
    k.w[1] = (float)w1d;
    k.w[2] = (float)w2d;
    const double wd[3] = {4.0 / 9.0, 1.0 / 9.0, 1.0 / 36.0};
    const double bbp_h = 1.5 * C.dt / (2.0 * C.taum + C.dt), cbp_h = C.taum * bbp_h;
    const double bh_h = C.hdt / (2.0 * C.taum + C.hdt), ch_h = C.taum * bh_h;
    // AC-p: tau = Mu/(rho RT) with Mu = rho RT Tau0 (D2Q9AC.cpp:34-60) -> Tau0
    const double bbp0 = 1.5 * C.dt / (2.0 * C.tau0 + C.dt), cbp0 = C.tau0 * bbp0;
    const double bh0 = C.hdt / (2.0 * C.tau0 + C.hdt), ch0 = C.tau0 * bh0;
    float fb_bp_h, fb_h, fb_bp0, fb0;
    split1((float)(1.0 - bbp_h), k.omb_bp_h, fb_bp_h);
    split1((float)(1.0 - bh_h), k.omb_h, fb_h);
    split1((float)(1.0 - bbp0), k.omb_bp0, fb_bp0);
    split1((float)(1.0 - bh0), k.omb0, fb0);
    const double b_bp_h = fb_bp_h, b_h = fb_h, b_bp0 = fb_bp0, b0 = fb0;
    for (int t = 0; t < 3; ++t) {
        k.bpw_h[t] = (float)(b_bp_h * k.w[t]);
        k.bw_h[t] = (float)(b_h * k.w[t]);
        k.bpw0[t] = (float)(b_bp0 * k.w[t]);
        k.bw0[t] = (float)(b0 * k.w[t]);
        k.cbpw_h[t] = (float)(cbp_h * wd[t]);
        k.chw_h[t] = (float)(ch_h * wd[t]);
        k.cbpw0[t] = (float)(cbp0 * wd[t]);
        k.chw0[t] = (float)(ch0 * wd[t]);
    }
    k.bprest0 = (float)(b_bp0 * ((double)k.w[0] - 1.0) / C.rt);
    k.hneq = (float)(2.0 * C.taum / (2.0 * C.taum + C.dt));
    k.fneq0 = (float)(2.0 * C.tau0 / (2.0 * C.tau0 + C.dt));
    k.hdx = (float)(0.5 * C.dx);
    k.uw_n = (float)((0.5 - A) / 12.0);
    k.uw_t = (float)(A / 12.0);
    k.uw_0 = (float)(1.0 / 24.0);
    return k;
}

// WALLS template argument: 0 periodic y, 1 bounce-back, 2 non-equilibrium extrapolation
int walls_code(const Consts &C) { return C.walls ? (C.nee ? 2 : 1) : 0; }

}  // namespace

struct Batch32Plan {
    BatchCase *d_cases = nullptr;
    int2 *d_work = nullptr;
    int ncases = 0, nwork = 0, model = 0, walls = 0, tile = MPD_TILE_16X15;
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
int batch32_plan(Batch32Plan *P, const BatchMember *m, int n, int tile, hipStream_t st)
{
    const int tx = tile == MPD_TILE_32X16 ? 32 : 16, ty = tile == MPD_TILE_16X15 ? 15 : 16;
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
        const int ntiles = ((C.nx + tx - 1) / tx) * ((C.ny + ty - 1) / ty);
        for (int t = 0; t < ntiles; ++t) list.push_back(make_int2(c, t));
    }
    const int per = ((int)list.size() + 7) / 8;
    std::vector<int2> work((size_t)per * 8, make_int2(-1, 0));
    for (int b = 0; b < per * 8; ++b) {
        const int p = (b % 8) * per + b / 8;
        if (p < (int)list.size()) work[(size_t)b] = list[(size_t)p];
    }
    // growing a table frees the old one first; a failed allocation leaves the
    // table empty (capacity 0), never a freed pointer
    if ((size_t)n > P->cap_cases) {
        if (P->d_cases) (void)hipFree(P->d_cases);
        P->d_cases = nullptr;
        P->cap_cases = 0;
        if (hipMalloc(&P->d_cases, sizeof(BatchCase) * (size_t)n) != hipSuccess) {
            P->d_cases = nullptr;
            return -1;
        }
        P->cap_cases = (size_t)n;
    }
    if (work.size() > P->cap_work) {
        if (P->d_work) (void)hipFree(P->d_work);
        P->d_work = nullptr;
        P->cap_work = 0;
        if (hipMalloc(&P->d_work, sizeof(int2) * work.size()) != hipSuccess) {
            P->d_work = nullptr;
            return -1;
        }
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
    P->walls = n ? walls_code(*m[0].C) : 0;
    P->tile = tile == MPD_TILE_32X16 || tile == MPD_TILE_16X16 ? tile : MPD_TILE_16X15;
    return 0;
}

template <int MODEL, int TX, int TY>
void batch32_launch(const Batch32Plan *P, int parity, hipStream_t st)
{
    dim3 g((unsigned)P->nwork), b(TileK<TX, TY>::NT);
    if (P->walls == 2)
        hipLaunchKernelGGL((k_fused32_batch<MODEL, 2, TX, TY>), g, b, 0, st, P->d_cases, P->d_work, P->nwork, parity);
    else if (P->walls)
        hipLaunchKernelGGL((k_fused32_batch<MODEL, 1, TX, TY>), g, b, 0, st, P->d_cases, P->d_work, P->nwork, parity);
    else
        hipLaunchKernelGGL((k_fused32_batch<MODEL, 0, TX, TY>), g, b, 0, st, P->d_cases, P->d_work, P->nwork, parity);
}

template <int MODEL>
void batch32_launch_shape(const Batch32Plan *P, int parity, hipStream_t st)
{
    if (P->tile == MPD_TILE_32X16) batch32_launch<MODEL, 32, 16>(P, parity, st);
    else if (P->tile == MPD_TILE_16X16) batch32_launch<MODEL, 16, 16>(P, parity, st);
    else batch32_launch<MODEL, 16, 15>(P, parity, st);
}

void batch32_step(const Batch32Plan *P, int parity, hipStream_t st)
{
    if (!P->nwork) return;
    if (P->model == MPD_MODEL_AC_HCZ) batch32_launch_shape<MPD_MODEL_AC_HCZ>(P, parity, st);
    else batch32_launch_shape<MPD_MODEL_AC_P>(P, parity, st);
}

template <int MODEL, int TX, int TY, int SC = 0>
void fused32_launch(const K32 &k, const FusedPlanes32 &in, const FusedPlanes32 &out, const Consts &C, int64_t N,
                    hipStream_t st)
{
    const int ntx = (C.nx + TX - 1) / TX, nty = (C.ny + TY - 1) / TY;
    const int ntiles = ntx * nty;
    dim3 g((unsigned)(((ntiles + 7) / 8) * 8)), b(TileK<TX, TY>::NT);
    const int w = walls_code(C);
    if (w == 2)
        hipLaunchKernelGGL((k_fused32<MODEL, 2, TX, TY, MPD_FLUX_CD, SC>), g, b, 0, st, k, C.nx, C.ny, in, out, N, ntx,
                           ntiles);
    else if (w)
        hipLaunchKernelGGL((k_fused32<MODEL, 1, TX, TY, MPD_FLUX_CD, SC>), g, b, 0, st, k, C.nx, C.ny, in, out, N, ntx,
                           ntiles);
    else
        hipLaunchKernelGGL((k_fused32<MODEL, 0, TX, TY, MPD_FLUX_CD, SC>), g, b, 0, st, k, C.nx, C.ny, in, out, N, ntx,
                           ntiles);
}

template <int MODEL>
void fused32_launch_shape(const K32 &k, const FusedPlanes32 &in, const FusedPlanes32 &out, const Consts &C,
                          int64_t N, int tile, bool one_round, hipStream_t st)
{
    if (tile == MPD_TILE_32X16 && one_round) fused32_launch<MODEL, 32, 16, ST_WT>(k, in, out, C, N, st);
    else if (tile == MPD_TILE_32X16) fused32_launch<MODEL, 32, 16>(k, in, out, C, N, st);
    else if (tile == MPD_TILE_16X16) fused32_launch<MODEL, 16, 16>(k, in, out, C, N, st);
    else fused32_launch<MODEL, 16, 15>(k, in, out, C, N, st);
}

int fused32_tile_auto(int nx, int ny, int cus, int64_t state_bytes)
{
    // One round: a launch whose tiles all fit the resident workgroups (4 per
    // CU of 256 threads, 2 of 512) has no partly filled last round: 32x16
    // where its tiles fit one round and 16x15 tiles would not, else 16x16
    // under the same condition.  Several rounds: 16x15 (the fewest
    // instructions per cell) while the two state sets fit the 256 MiB
    // Infinity Cache; beyond it the step streams from HBM and 32x16, with
    // fewer halo lines re-read per cell, is faster.  Measured on MI355X (256
    // CUs), FP32 HCZ with walls, us per step 16x15 / 16x16 / 32x16: 256x1024
    // 19.2 / 18.1 / 17.5, 1024x256 20.5 / 19.0 / 17.2, 512x512 19.5 / 18.2 /
    // 18.0, 128x128 8.9 / 9.5 / 10.7, 768x768 32.8 / 35.5 / 37.2, 1024x1024
    // 51.6 / 53.7 / 54.2, 1280x1280 118 / 117 / 104, 2048x1024 137 / 132 /
    // 121, 2048x2048 264 / 252 / 227, 3072x2048 415 / 410 / 350.
    const int64_t slots = 4 * (int64_t)(cus > 0 ? cus : 256);
    const int64_t t15 = (int64_t)((nx + 15) / 16) * ((ny + 14) / 15);
    const int64_t t16 = (int64_t)((nx + 15) / 16) * ((ny + 15) / 16);
    const int64_t t32 = (int64_t)((nx + 31) / 32) * ((ny + 15) / 16);
    const int64_t mall = (int64_t)256 << 20;
    if (t15 <= slots) return MPD_TILE_16X15;
    if (t32 <= slots / 2) return MPD_TILE_32X16;
    if (t16 <= slots) return MPD_TILE_16X16;
    return state_bytes > mall ? MPD_TILE_32X16 : MPD_TILE_16X15;
}

template <int MODEL>
void fused32_launch_uw(const K32 &k, const FusedPlanes32 &in, const FusedPlanes32 &out, const Consts &C, int64_t N,
                       hipStream_t st)
{
    const int ntx = (C.nx + 15) / 16, nty = (C.ny + 14) / 15;
    const int ntiles = ntx * nty;
    dim3 g((unsigned)(((ntiles + 7) / 8) * 8)), b(256);
    const int w = walls_code(C);
    if (w == 2)
        hipLaunchKernelGGL((k_fused32<MODEL, 2, 16, 15, MPD_FLUX_UW>), g, b, 0, st, k, C.nx, C.ny, in, out, N, ntx,
                           ntiles);
    else if (w)
        hipLaunchKernelGGL((k_fused32<MODEL, 1, 16, 15, MPD_FLUX_UW>), g, b, 0, st, k, C.nx, C.ny, in, out, N, ntx,
                           ntiles);
    else
        hipLaunchKernelGGL((k_fused32<MODEL, 0, 16, 15, MPD_FLUX_UW>), g, b, 0, st, k, C.nx, C.ny, in, out, N, ntx,
                           ntiles);
}

void fused32_step(const FusedPlanes32 &in, const FusedPlanes32 &out, const Consts &C, int64_t N, int tile,
                  int cus, hipStream_t st)
{
    // one round: every 32x16 tile resident at once (two 512-thread workgroups per CU)
    const bool one_round = (int64_t)((C.nx + 31) / 32) * ((C.ny + 15) / 16) <= 2 * (int64_t)(cus > 0 ? cus : 256);
    const K32 k = make_k32(C);
    if (C.flux == MPD_FLUX_UW) {
        // UW faces: 16x15 tiles (the 3-cell halo)
        if (C.model == MPD_MODEL_AC_HCZ) fused32_launch_uw<MPD_MODEL_AC_HCZ>(k, in, out, C, N, st);
        else fused32_launch_uw<MPD_MODEL_AC_P>(k, in, out, C, N, st);
        return;
    }
    if (C.model == MPD_MODEL_AC_HCZ) fused32_launch_shape<MPD_MODEL_AC_HCZ>(k, in, out, C, N, tile, one_round, st);
    else fused32_launch_shape<MPD_MODEL_AC_P>(k, in, out, C, N, tile, one_round, st);
}

}  // namespace mpd
