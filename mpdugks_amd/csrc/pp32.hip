// FP32-bulk pseudopotential step: D2Q9 (CD and UW faces) and D2V16 (CD faces),
// periodic mesh.
//
// The phases of the reference step for the pseudopotential models
// (Ark/Cpp/DUGKS2DSolver.cpp:164-328 with Ark/Lib/D2Q9PP.cpp, Ark/Lib/D2V16.cpp,
// Ark/Cpp/PseudoP.cpp): bar-plus with the cell force, face values, face macros
// with the face's previous force, face Psi, face force, half-step BGK and
// micro-flux, cell update, cell Psi.  Populations, bar-plus, face values, face
// macros and fluxes are FP32.  The Carnahan-Starling pressure is FP64 and so is
// the difference p - rho RT it enters Psi by (near the spinodal the two are
// nearly equal); Psi is rounded to FP32 once.
//
// State set (fused.h, PP32Planes): f~[q], U, V, the face Psi of the step that
// produced the set signed by scG (v = scG Psi; x-faces, then y-faces; face t of
// a direction is the one west / south of cell t), and the density in FP64.
// Everything else the reference keeps between steps is recomputed from it:
//  - cell Psi and scG: the FP64 EOS of the density;
//  - the cell force -scG Psi grad(Psi) (Update_PseudoForce(cell), PseudoP.cpp:31-38):
//    the 9-point gradient of the cell Psi;
//  - a face's previous force (D2Q9PP.cpp:117-118 keeps it per face):
//    -v grad|v| over the face's own lattice (GradSchemeBasic.h:181-208;
//    PseudoP.cpp:39-46) of the previous step's face Psi -- the values the
//    previous step used, so the force is the same number.
// The first step after a conversion from the explicit state takes the forces
// of that state instead (the shared planes), as they need not derive from Psi.
//
// Mass.  The density is the conserved variable, kept in FP64: it moves only by
// the face mass fluxes (each face's sum over populations, shared by both cells
// with opposite signs).  The populations follow it: D2Q9's rest population
// takes the remainder; D2V16, which has none, spreads it over the four
// populations of the diagonal c shell, whose velocities sum to zero, so the
// momentum is untouched.
//
// Compiled without floating-point contraction, the multiply-adds written out:
// a face on a tile edge is computed by both tiles that use it, in different
// inlined copies of the face code, and must give the same bits in both (so the
// step does not depend on the tile shape and a face's mass flux leaves one cell
// exactly as it enters the other).
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "d2v16.h"
#include "fused.h"
#include "physics.h"

namespace mpd {

namespace {

// ---------------- velocity sets ----------------
// Directions as integer multiples of the lattice speed ms (MaSpan), opposite
// pairs (q, opp(q)) with pair p's first member first(p), and the population
// set that takes the density remainder.
struct L9 {
    static constexpr int Q = 9, NP = 4;
    static constexpr bool REST = true;
    static constexpr int ex(int q) { return ex9(q); }
    static constexpr int ey(int q) { return ey9(q); }
    static constexpr int first(int p) { return p + 1; }  // (1,5) (2,6) (3,7) (4,8)
    static constexpr int opp(int q) { return bb9(q); }
};
struct L16 {
    static constexpr int Q = 16, NP = 8;
    static constexpr bool REST = false;
    static constexpr int ex(int q) { return d2v16::ux(q) * (d2v16::shell(q) < 2 ? 1 : 2); }
    static constexpr int ey(int q) { return d2v16::uy(q) * (d2v16::shell(q) < 2 ? 1 : 2); }
    static constexpr int first(int p) { return (p >> 1) * 4 + (p & 1); }  // (0,2) (1,3) (4,6) (5,7) ...
    static constexpr int opp(int q) { return (q & ~3) | ((q + 2) & 3); }
};
template <class L, int DIR>
__host__ __device__ constexpr int en(int q) { return DIR == 0 ? L::ex(q) : L::ey(q); }
template <class L, int DIR>
__host__ __device__ constexpr int et(int q) { return DIR == 0 ? L::ey(q) : L::ex(q); }
// populations with a non-zero normal velocity on a face of direction DIR (the
// others carry no flux), in increasing q: count, z-th member, slot of q (or -1)
template <class L, int DIR>
__host__ __device__ constexpr int nm()
{
    int c = 0;
    for (int q = 0; q < L::Q; ++q) c += en<L, DIR>(q) != 0;
    return c;
}
template <class L, int DIR>
__host__ __device__ constexpr int mq(int z)
{
    for (int q = 0; q < L::Q; ++q)
        if (en<L, DIR>(q) != 0 && z-- == 0) return q;
    return -1;
}
template <class L, int DIR>
__host__ __device__ constexpr int slot(int q)
{
    int z = 0;
    for (int r = 0; r < q; ++r) z += en<L, DIR>(r) != 0;
    return en<L, DIR>(q) != 0 ? z : -1;
}
static_assert(nm<L9, 0>() == 6 && nm<L9, 1>() == 6 && nm<L16, 0>() == 12 && nm<L16, 1>() == 12, "flux sets");

// D2V16 equilibrium table in FP32 (the inverse moment matrix, d2v16.h)
struct A16 {
    float a[16][16];
};
constexpr A16 make_a16()
{
    A16 t{};
    for (int j = 0; j < 16; ++j)
        for (int m = 0; m < 16; ++m) t.a[j][m] = (float)d2v16::inv(j, m);
    return t;
}
constexpr A16 kA16 = make_a16();

// FP32 constants of the step, derived on the host in FP64 from Consts: every
// division by a constant of the reference's expressions is a product with its
// reciprocal, and the lattice velocities fold at compile time.
struct KPP {
    float ms;              // lattice speed (MaSpan; 1.2 for D2V16)
    float ms_rt, inv_rt, rt, lambda0, hdt, half_hdt, bdt;  // bdt = dt + hdt
    float dt;
    float ra[2], rb[2];    // CD face value for e_n = +1, +2: ra o + rb n (ra = 1/2 + e_n A, A = hdt ms / dx)
    float rt4;             // tangential factor per unit e_t: hdt ms / (2 twodx)
    float inv_d2x6;        // 9-point gradient
    float w[3];            // D2Q9 weights: rest, axis, diagonal
    float wirt[3];         // w / RT
    float omb_u, bbp_u, cbp_u;  // bar-plus factors of a uniform relaxation time (tau_u > 0): 1 - b, b, c
    float tau_u;           // the uniform cell relaxation time, or 0
    float omb_h, bh, ch;   // face BGK: (1 - bh) f + bh f_eq + ch f_So
    // (1 - b, b) pairs sum to exactly 1 (split1); the equilibria are computed
    // for the density b rho, which yields b f_eq
    float dtv, flux;       // dt / V; dx ms (micro-flux per unit e_n)
    float uw_n, uw_t, uw_0;  // UW face value, 6-point gradient sums (x 12 dx)
    double pp_k, t_cs, drt;  // FP64 EOS: K, T_CS, RT
};

KPP make_kpp(const Consts &C, int q, double tau_u)
{
    KPP k;
    const double ms = q == 16 ? d2v16::C : C.xi[0][1];
    k.ms = (float)ms;
    k.ms_rt = (float)(ms / C.rt);
    k.inv_rt = (float)(1.0 / C.rt);
    k.rt = (float)C.rt;
    k.lambda0 = (float)C.lambda0;
    k.hdt = (float)C.hdt;
    k.half_hdt = (float)(0.5 * C.hdt);
    k.bdt = (float)(C.dt + C.hdt);
    k.dt = (float)C.dt;
    const double A = C.hdt * ms / C.dx;
    for (int e = 0; e < 2; ++e) {
        k.ra[e] = (float)(0.5 + (e + 1) * A);
        k.rb[e] = (float)(0.5 - (e + 1) * A);
    }
    k.rt4 = (float)(C.hdt * ms / (2.0 * C.twodx));
    k.inv_d2x6 = (float)(1.0 / C.d2x6);
    k.w[0] = (float)C.w[0];
    k.w[1] = (float)C.w[1];
    k.w[2] = (float)C.w[2];
    for (int c = 0; c < 3; ++c) k.wirt[c] = (float)(C.w[c] / C.rt);
    // Cell_2D.cpp:120-125: bBP = 1 - (2 tau - hdt)/(2 tau + dt) = (dt + hdt)/(2 tau + dt), cBP = tau bBP
    k.tau_u = tau_u > 0 ? (float)tau_u : 0.f;
    k.omb_u = k.bbp_u = 0.f;
    if (tau_u > 0) split1((float)(1.0 - (C.dt + C.hdt) / (2.0 * tau_u + C.dt)), k.omb_u, k.bbp_u);
    k.cbp_u = tau_u > 0 ? (float)(tau_u * (C.dt + C.hdt) / (2.0 * tau_u + C.dt)) : 0.f;
    split1((float)(1.0 - C.bh_pp), k.omb_h, k.bh);
    k.ch = (float)C.ch_pp;
    k.dtv = (float)C.dtv;
    k.flux = (float)(C.dx * ms);
    k.uw_n = (float)((0.5 - A) / 12.0);
    k.uw_t = (float)(A / 12.0);
    k.uw_0 = (float)(1.0 / 24.0);
    k.pp_k = C.pp_k;
    k.t_cs = C.t_cs;
    k.drt = C.rt;
    return k;
}

// weight class of D2Q9 population q: 0 rest, 1 axis (odd q), 2 diagonal (even q)
__host__ __device__ constexpr int wcls(int q) { return q == 0 ? 0 : ((q & 1) ? 1 : 2); }

// ---------------- FP64 equation of state ----------------
// scG Psi of a density (MacroQuantity::calcPsi, Ark/Header/Macroscopic.h:64-68,
// with the Carnahan-Starling pressure, Ark/Lib/nonidealEoS.h:59-73, a = 1,
// b = 4): d = p - rho RT in FP64, Psi = sqrt(2 |d|) (dividing by scG = +-1),
// scG = sign(d); NaN where d = 0 (the reference's 0/0).  The division by
// (1 - rho)^3 is the hardware reciprocal (relative error below 2^-22) refined by
// one Newton step (below 2^-44): d keeps ~1e-13 of p, far below the FP32
// rounding of Psi however close p and rho RT are.
__device__ __forceinline__ float pp_vsig(const KPP &k, double rho)
{
    const double om = 1.0 - rho;
    const double den = om * om * om;
    double r = __builtin_amdgcn_rcp(den);
    r = fma(r, fma(-den, r, 1.0), r);
    const double num = fma(rho, fma(rho, 1.0 - rho, 1.0), 1.0);  // 1 + rho + rho^2 - rho^3
    const double p = k.pp_k * fma(rho * k.t_cs * num, r, -(rho * rho));
    const double d = fma(-rho, k.drt, p);
    const float psi = sqrtf((float)fabs(2.0 * d));
    return d > 0.0 ? psi : (d < 0.0 ? -psi : __builtin_nanf(""));
}

// 9-point isotropic gradient (update_MQ_x/y, Ark/Lib/GradSchemeBasic.h:153-208)
// of |v| around a site, times -v of the site: the pseudopotential force
// -scG Psi grad(Psi) (Update_PseudoForce, Ark/Cpp/PseudoP.cpp:31-46).  e, n, w, s
// and the diagonals are the neighbours' signed Psi.
__device__ __forceinline__ void pp_force(const KPP &k, float v, float e, float n, float w, float s, float ne, float nw,
                                         float sw, float se, float &Fx, float &Fy)
{
    e = fabsf(e); n = fabsf(n); w = fabsf(w); s = fabsf(s);
    ne = fabsf(ne); nw = fabsf(nw); sw = fabsf(sw); se = fabsf(se);
    const float gx = fmaf(4.f, e - w, (ne - nw) + (se - sw)) * k.inv_d2x6;
    const float gy = fmaf(4.f, n - s, (ne - se) + (nw - sw)) * k.inv_d2x6;
    Fx = gx * -v;
    Fy = gy * -v;
}

// f(std::integral_constant<int, i>) for i = 0 .. N-1: loop indices that are
// compile-time constants inside the body
template <int N, class F>
__device__ __forceinline__ void unroll(F &&f)
{
    if constexpr (N > 0) {
        unroll<N - 1>(f);
        f(std::integral_constant<int, N - 1>{});
    }
}

// ---------------- equilibrium and source ----------------
// f_eq and the He-Shan-Doolen source f_So = F.(xi - u) f_eq / (rho RT) of the
// populations of the pairs PMASK(p) selects (and the rest population, REST)
// (Update_phi_Eq / Update_DVDF_Source: D2Q9PP.cpp:51-62,137-145; D2V16.cpp:157-166
// with Macro16 at the reference temperature, :125-145).  Every equilibrium is
// rho times a function of u (p = rho RT), so f_So = F.(xi - u) (f_eq / rho) / RT
// takes no division by rho.  D2Q9 pairs (q, q') share
// B = 1 + u1^2/2 - |u|^2 Lambda0: f_eq = w rho (B +- u1), u1 = xi_q.u / RT.
struct Site {
    float rho, U, V, Fx, Fy;
};
template <int A, int B>
__device__ __forceinline__ float edot(float x, float y)
{
    if (A == 0 && B == 0) return 0.f;
    if (A == 0) return B > 0 ? y : -y;
    if (B == 0) return A > 0 ? x : -x;
    if (A > 0) return B > 0 ? x + y : x - y;
    return B > 0 ? y - x : -(x + y);
}
template <class L, bool REST, class PM>
__device__ __forceinline__ void feq_src(const KPP &k, const Site &s, float *feq, float *fso, PM pmask)
{
    const float uF = fmaf(s.U, s.Fx, s.V * s.Fy);
    if constexpr (L::Q == 9) {
        const float c0 = fmaf(-k.lambda0, fmaf(s.U, s.U, s.V * s.V), 1.f);
        if (REST) {
            feq[0] = (k.w[0] * s.rho) * c0;
            fso[0] = (-uF * k.wirt[0]) * c0;
        }
        unroll<L::NP>([&](auto P) {
            constexpr int p = decltype(P)::value, q = L::first(p), qp = L::opp(q);
            constexpr int EX = L::ex(q), EY = L::ey(q), c = wcls(q);
            if (!pmask(p)) return;
            const float u1 = k.ms_rt * edot<EX, EY>(s.U, s.V);
            const float eF = edot<EX, EY>(s.Fx, s.Fy);
            const float B = fmaf(0.5f * u1, u1, c0);
            const float bp = B + u1, bm = B - u1, wr = k.w[c] * s.rho;
            feq[q] = wr * bp;
            feq[qp] = wr * bm;
            fso[q] = (fmaf(k.ms, eF, -uF) * k.wirt[c]) * bp;
            fso[qp] = (fmaf(-k.ms, eF, -uF) * k.wirt[c]) * bm;
        });
    } else {
        float Mo[16];
        const float rt = k.rt;
        d2v16::macro16<float>(Mo, 1.f, s.U, s.V, rt, rt);
        unroll<L::NP>([&](auto P) {
            constexpr int p = decltype(P)::value;
            if (!pmask(p)) return;
            unroll<2>([&](auto S) {
                constexpr int q = decltype(S)::value == 0 ? L::first(p) : L::opp(L::first(p));
                float a = 0.f;
                unroll<16>([&](auto MM) {
                    constexpr int m = decltype(MM)::value;
                    if constexpr (kA16.a[q][m] != 0.f) a = fmaf(kA16.a[q][m], Mo[m], a);
                });
                const float eF = fmaf((float)L::ex(q), s.Fx, (float)L::ey(q) * s.Fy);
                feq[q] = s.rho * a;
                fso[q] = (fmaf(k.ms, eF, -uF) * k.inv_rt) * a;
            });
        });
    }
}

// ---------------- global memory ----------------
// Buffer resources over one state set's float block, its FP64 density and the
// shared planes: a per-lane 32-bit byte offset plus a uniform plane offset
// (the host refuses meshes whose planes would span 4 GiB).
using rsrc_t = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ rsrc_t make_rsrc(const void *p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float ldf(rsrc_t r, uint32_t n, uint32_t plane_bytes)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, n * 4u, plane_bytes, 0));
}
__device__ __forceinline__ void stf(rsrc_t r, uint32_t n, float v, uint32_t plane_bytes)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, n * 4u, plane_bytes, 0);
}
__device__ __forceinline__ double ldd(rsrc_t r, uint32_t n)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, n * 8u, 0, 0));
}
__device__ __forceinline__ void std_(rsrc_t r, uint32_t n, double v)
{
    using u2 = unsigned int __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, n * 8u, 0, 0);
}
// shared planes: tau, Fx, Fy [N], ffx, ffy [2N], consecutive (pp32_alloc)
enum : uint32_t { SH_TAU = 0, SH_FX = 1, SH_FY = 2, SH_FFX = 3, SH_FFY = 5, SH_PLANES = 7 };

struct Rs {
    rsrc_t f, rho, sh;
    uint32_t pb;  // bytes per cell plane
};
__device__ __forceinline__ Rs make_rs(const float *f, const double *rho, const float *sh, int64_t N, int q)
{
    Rs r;
    r.pb = (uint32_t)(N * 4);
    r.f = make_rsrc(f, (uint32_t)(q + 4) * r.pb);
    r.rho = make_rsrc(rho, (uint32_t)(N * 8));
    r.sh = make_rsrc(sh, SH_PLANES * r.pb);
    return r;
}

// periodic wrap; `small` (uniform) selects the general form for meshes narrower
// than a tile's regions, otherwise one conditional add or subtract suffices
__device__ __forceinline__ int wrap(int i, int n, bool small)
{
    if (small) {
        const int r = i % n;
        return r < 0 ? r + n : r;
    }
    i += i < 0 ? n : 0;
    return i >= n ? i - n : i;
}

__device__ __forceinline__ int tile_of_block(int b, int ntiles)
{
    const int per = (ntiles + 7) / 8;
    const int t = (b % 8) * per + b / 8;  // XCD k (block b mod 8) takes a contiguous band of tiles
    return t < ntiles ? t : -1;
}

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }


// ---------------- tile geometry ----------------
// One TX x TY tile per workgroup (NT threads, one tile cell per thread in the
// update).  Regions around the tile, relative to its first cell:
//  bar-plus cells -RO .. T+RO-1 (the cells the face values read; RO = 2 for CD,
//    3 for UW, whose 6-point gradients reach one cell further);
//  Psi cells -PO .. T+PO-1, PO = RO + 1 (the cell force of every bar-plus cell);
//  x-faces fx -1..TX+1, fy -1..TY and y-faces fx -1..TX, fy -1..TY+1: the
//    faces whose Psi the force of a tile face reads (and whose previous Psi
//    the previous force of a tile face reads);
//  tile faces: x-faces fx 0..TX, fy 0..TY-1; y-faces fx 0..TX-1, fy 0..TY.
// Face fx of a row is the face west of cell fx (y-faces: south), so face (fx,
// fy) has the index of cell (fx, fy): the face-Psi arrays use the Psi region's
// layout, one per direction.
template <class L, int TX_, int TY_, int H_>
struct PTile {
    static constexpr int TX = TX_, TY = TY_, H = H_;
    static constexpr int NT = TX * TY > 256 ? 512 : 256;
    static constexpr int NIN = TX * TY;
    static constexpr int RO = 1 + H, PO = RO + 1;
    static constexpr int RX = TX + 2 * RO, RY = TY + 2 * RO, RN = RX * RY;
    static constexpr int QX = TX + 2 * PO, QY = TY + 2 * PO, QN = QX * QY;
    static constexpr int NXF = (TX + 1) * TY, NYF = TX * (TY + 1), FN = NXF + NYF;
    static constexpr int NXR = 2 * (TX + 3) + 2 * TY, NYR = 2 * (TY + 3) + 2 * TX;  // Psi-only faces
    static constexpr int NXA = NXF < NT ? NXF : NT, NYA = NYF < NT ? NYF : NT;  // first two face passes
    static constexpr int EXF = NXF - NXA, EYF = NYF - NYA;                  // tile faces of the third pass
    // third pass: x-faces on threads [0, EXF + NXR), y-faces from the next wave
    static constexpr int YC0 = (EXF + NXR + 63) / 64 * 64, NC = YC0 + EYF + NYR;
    static constexpr int NM = nm<L, 0>();                                     // flux populations per face
    // LDS, floats: a union of the stages up to the faces (bar-plus pairs, rest,
    // cell Psi, cell density) and of the flux stage (NM fluxes + mass per tile face)
    static constexpr int U1 = 2 * L::NP * RN + (L::REST ? RN : 0) + 2 * QN;
    static constexpr int U2 = (NM + 1) * FN;
    static constexpr int NU = U1 > U2 ? U1 : U2;
    static_assert(NIN <= NT && NC <= NT, "face passes");
    static_assert(PO >= 2, "face-Psi regions inside the Psi region");
};

// position (cx, cy) of ring item r of a region of halo HW around the tile
// (width TX + 2 HW): the HW bottom rows, the HW top rows, then the 2 HW halo
// cells of each tile row
template <int TX, int TY, int HW>
__device__ __forceinline__ void ring_pos(int r, int &cx, int &cy)
{
    constexpr int W = TX + 2 * HW;
    if (r < HW * W) {
        cy = r / W - HW;
        cx = r % W - HW;
    } else if (r < 2 * HW * W) {
        r -= HW * W;
        cy = TY + r / W;
        cx = r % W - HW;
    } else {
        r -= 2 * HW * W;
        const int c = r % (2 * HW);
        cy = r / (2 * HW);
        cx = c < HW ? c - HW : TX + (c - HW);
    }
}

// Psi-only face r of a direction (the faces of the face-Psi region that are
// not tile faces).  x-faces: rows fy = -1 and fy = TY (fx -1..TX+1), then
// columns fx = -1 and fx = TX+1 (fy 0..TY-1); y-faces: columns fx = -1 and
// fx = TX (fy -1..TY+1), then rows fy = -1 and fy = TY+1 (fx 0..TX-1)
template <class G, int DIR>
__device__ __forceinline__ void ring_face(int r, int &fx, int &fy)
{
    constexpr int TX = G::TX, TY = G::TY;
    if (DIR == 0) {
        constexpr int W = TX + 3;
        if (r < 2 * W) { fy = r < W ? -1 : TY; fx = (r < W ? r : r - W) - 1; }
        else { r -= 2 * W; fx = r < TY ? -1 : TX + 1; fy = r < TY ? r : r - TY; }
    } else {
        constexpr int W = TY + 3;
        if (r < 2 * W) { fx = r < W ? -1 : TX; fy = (r < W ? r : r - W) - 1; }
        else { r -= 2 * W; fy = r < TX ? -1 : TY + 1; fx = r < TX ? r : r - TX; }
    }
}

// index of a face (or a cell) in the Psi-region layout
template <class G>
__device__ __forceinline__ int qidx(int cx, int cy)
{
    return (cy + G::PO) * G::QX + (cx + G::PO);
}

// ---------------- face values ----------------
// CD face values (CD_Interior_phi_Bh, DUGKS2DSolver.cpp:604-687) of the pair
// (q, opp(q)) from the bar-plus pair plane b: 1/2 (o + n) - hdt xi_n (n - o)/dx
// - hdt xi_t (1/2 (tn_hi - tn_lo) + 1/2 (to_hi - to_lo))/twodx, with o the
// owner (west / south) and n the neighbour cell.  o, n, op, om, np, nm: region
// indices of the owner, the neighbour, and their tangential neighbours (+/-).
struct FaceIdx {
    int o, n, op, om, np, nm;
};
template <class G, int DIR>
__device__ __forceinline__ FaceIdx face_idx(int fx, int fy)
{
    constexpr int RX = G::RX, RO = G::RO;
    const int ox = DIR == 0 ? fx - 1 : fx, oy = DIR == 0 ? fy : fy - 1;
    const int o = (oy + RO) * RX + (ox + RO);
    const int dn = DIR == 0 ? 1 : RX, dt = DIR == 0 ? RX : 1;
    return FaceIdx{o, o + dn, o + dt, o - dt, o + dn + dt, o + dn - dt};
}
template <class L, int DIR, int P>
__device__ __forceinline__ float2 recon(const KPP &k, const float2 *b, const FaceIdx &x)
{
    constexpr int q = L::first(P);
    constexpr int EN = en<L, DIR>(q), ET = et<L, DIR>(q);
    constexpr int AN = EN < 0 ? -EN : EN, AT = ET < 0 ? -ET : ET;
    const float2 bo = b[x.o], bn = b[x.n];
    float2 v;
    if constexpr (EN == 0) {
        v.x = 0.5f * (bo.x + bn.x);
        v.y = 0.5f * (bo.y + bn.y);
    } else {
        const float ra = k.ra[AN - 1], rb = k.rb[AN - 1];
        // q has e_n = EN, its partner -EN
        v.x = EN > 0 ? fmaf(rb, bn.x, ra * bo.x) : fmaf(ra, bn.x, rb * bo.x);
        v.y = EN > 0 ? fmaf(ra, bn.y, rb * bo.y) : fmaf(rb, bn.y, ra * bo.y);
    }
    if constexpr (ET != 0) {
        const float2 np = b[x.np], nm = b[x.nm], op = b[x.op], om = b[x.om];
        const float c = (ET > 0 ? -k.rt4 : k.rt4) * (float)AT;
        v.x = fmaf(c, (np.x - nm.x) + (op.x - om.x), v.x);
        v.y = fmaf(-c, (np.y - nm.y) + (op.y - om.y), v.y);
    }
    return v;
}
// rest population (e = 0): the mean of both cells
__device__ __forceinline__ float recon_rest(const float *b, const FaceIdx &x) { return 0.5f * (b[x.o] + b[x.n]); }

// UW face values (UW_Interior_phi_Bh + Update_phi_fBh, DUGKS2DSolver.cpp:560-603,
// 688-706; gradients Grad_VS_6points, GradSchemeBasic.h:95-126) of the pair:
// the upwind cell's value plus its 6-point gradient along x_f - h xi - x_c; for
// e_n = 0 the mean of both cells'.  Gradient sums g = 12 dx grad.
template <class G>
__device__ __forceinline__ void grad6(const float *b, int c, float &gx, float &gy)
{
    constexpr int RX = G::RX;
    const float e = b[c + 1], w = b[c - 1], n = b[c + RX], s = b[c - RX];
    const float ne = b[c + RX + 1], nw = b[c + RX - 1], sw = b[c - RX - 1], se = b[c - RX + 1];
    gx = fmaf(4.f, e - w, (ne - nw) + (se - sw));
    gy = fmaf(4.f, n - s, (ne - se) + (nw - sw));
}
template <class G>
__device__ __forceinline__ void grad6(const float2 *b, int c, float2 &gx, float2 &gy)
{
    constexpr int RX = G::RX;
    const float2 e = b[c + 1], w = b[c - 1], n = b[c + RX], s = b[c - RX];
    const float2 ne = b[c + RX + 1], nw = b[c + RX - 1], sw = b[c - RX - 1], se = b[c - RX + 1];
    gx.x = fmaf(4.f, e.x - w.x, (ne.x - nw.x) + (se.x - sw.x));
    gx.y = fmaf(4.f, e.y - w.y, (ne.y - nw.y) + (se.y - sw.y));
    gy.x = fmaf(4.f, n.x - s.x, (ne.x - se.x) + (nw.x - sw.x));
    gy.y = fmaf(4.f, n.y - s.y, (ne.y - se.y) + (nw.y - sw.y));
}
// one side's value: the cell's bar-plus plus its normal (cn) and tangential
// (e_t) gradient terms
__device__ __forceinline__ float uw_one(const KPP &k, int e_t, float bc, float gn, float gt, float cn)
{
    float v = fmaf(cn, gn, bc);
    if (e_t != 0) v = fmaf(e_t > 0 ? -k.uw_t : k.uw_t, gt, v);
    return v;
}
template <class L, class G, int DIR, int P>
__device__ __forceinline__ float2 recon_uw(const KPP &k, const float2 *b, const FaceIdx &x)
{
    constexpr int q = L::first(P);
    constexpr int EN = en<L, DIR>(q), ET = et<L, DIR>(q);
    static_assert(EN >= -1 && EN <= 1 && ET >= -1 && ET <= 1, "UW faces: unit lattice velocities");
    float2 gxo, gyo, gxn, gyn;
    grad6<G>(b, x.o, gxo, gyo);
    grad6<G>(b, x.n, gxn, gyn);
    const float2 go_n = DIR == 0 ? gxo : gyo, go_t = DIR == 0 ? gyo : gxo;
    const float2 gn_n = DIR == 0 ? gxn : gyn, gn_t = DIR == 0 ? gyn : gxn;
    const float2 bo = b[x.o], bn = b[x.n];
    float2 v;
    if (EN == 0) {
        v.x = 0.5f * (uw_one(k, ET, bo.x, go_n.x, go_t.x, k.uw_0) +
                      uw_one(k, ET, bn.x, gn_n.x, gn_t.x, -k.uw_0));
        v.y = 0.5f * (uw_one(k, -ET, bo.y, go_n.y, go_t.y, k.uw_0) +
                      uw_one(k, -ET, bn.y, gn_n.y, gn_t.y, -k.uw_0));
    } else {
        // q (e_n = EN) is upwinded from the owner for EN > 0; its partner from the other side
        v.x = EN > 0 ? uw_one(k, ET, bo.x, go_n.x, go_t.x, k.uw_n)
                     : uw_one(k, ET, bn.x, gn_n.x, gn_t.x, -k.uw_n);
        v.y = EN > 0 ? uw_one(k, -ET, bn.y, gn_n.y, gn_t.y, -k.uw_n)
                     : uw_one(k, -ET, bo.y, go_n.y, go_t.y, k.uw_n);
    }
    return v;
}
template <class G>
__device__ __forceinline__ float recon_rest_uw(const KPP &k, const float *b, const FaceIdx &x, int dir)
{
    float gxo, gyo, gxn, gyn;
    grad6<G>(b, x.o, gxo, gyo);
    grad6<G>(b, x.n, gxn, gyn);
    return 0.5f * (fmaf(k.uw_0, dir == 0 ? gxo : gyo, b[x.o]) + fmaf(-k.uw_0, dir == 0 ? gxn : gyn, b[x.n]));
}

// all face values of a face of direction DIR
template <class L, class G, int DIR, int FLUX>
__device__ __forceinline__ void face_values(const KPP &k, const float2 *sp, const float *sr, const FaceIdx &x, float *fb)
{
    constexpr int RN = G::RN;
    if constexpr (L::REST) fb[0] = FLUX == MPD_FLUX_UW ? recon_rest_uw<G>(k, sr, x, DIR) : recon_rest(sr, x);
    unroll<L::NP>([&](auto P) {
        constexpr int p = decltype(P)::value;
        float2 v;
        if constexpr (FLUX == MPD_FLUX_UW) v = recon_uw<L, G, DIR, p>(k, sp + p * RN, x);
        else v = recon<L, DIR, p>(k, sp + p * RN, x);
        fb[L::first(p)] = v.x;
        fb[L::opp(L::first(p))] = v.y;
    });
}

// face density and velocity (Update_MacroVar_h, D2Q9PP.cpp:110-136; D2V16 with
// sixteen populations): rho = sum f, u = (ms sum e f + hdt/2 F_old) / rho
template <class L>
__device__ __forceinline__ float face_rho(const float *fb)
{
    float s = 0.f;
#pragma unroll
    for (int q = 0; q < L::Q; ++q) s += fb[q];
    return s;
}
template <class L>
__device__ __forceinline__ void moment1(const float *f, float &sx, float &sy)
{
    sx = 0.f;
    sy = 0.f;
#pragma unroll
    for (int q = 0; q < L::Q; ++q) {
        if (L::ex(q) == 1) sx += f[q];
        if (L::ex(q) == -1) sx -= f[q];
        if (L::ex(q) == 2) sx = fmaf(2.f, f[q], sx);
        if (L::ex(q) == -2) sx = fmaf(-2.f, f[q], sx);
        if (L::ey(q) == 1) sy += f[q];
        if (L::ey(q) == -1) sy -= f[q];
        if (L::ey(q) == 2) sy = fmaf(2.f, f[q], sy);
        if (L::ey(q) == -2) sy = fmaf(-2.f, f[q], sy);
    }
}

// ---------------- the one-launch tile kernel ----------------
//  S0: density (FP64) and signed Psi (FP64 EOS) of the Psi region, and the
//      previous face Psi of the face-Psi regions;
//  S1: cell force and bar-plus of the bar-plus region (tile cells first: each
//      keeps b + (b - f~)/3 and its force in registers for S4);
//  S2: face values, density and Psi of every face of the face-Psi regions; the
//      tile faces also their velocity (with the previous force) and the face
//      values that carry a flux, in registers;
//  S3: face force, half-step BGK (Update_phi_h, DUGKS2DSolver.cpp:402-423) and
//      micro-flux (:424-441) of the tile faces; their Psi is the new face state;
//  S4: Update_phi_T (:748-792) with the FP64 density, Update_MacroVar.
// FORCE: the forces derive from Psi (every step but the first after a
// conversion).  TAUU: every cell has the relaxation time k.tau_u.

template <class L, class G, int FLUX, bool FORCE, int DIR>
__device__ __forceinline__ void s2_face(const KPP &k, const Rs &in, const float2 *sp, const float *sr,
                                        const float *s_vold, float *s_v, int fx, int fy, bool tile_face, int i0,
                                        int j0, int nx, int ny, bool small, float *hv, float &hrho, float &hU,
                                        float &hV)
{
    constexpr int QX = G::QX;
    const FaceIdx x = face_idx<G, DIR>(fx, fy);
    float fb[L::Q];
    face_values<L, G, DIR, FLUX>(k, sp, sr, x, fb);
    const float rho = face_rho<L>(fb);
    const int c = qidx<G>(fx, fy);
    s_v[c] = pp_vsig(k, (double)rho);
    if (!tile_face) return;
    float Fx, Fy;
    if (FORCE) {
        const float *v = s_vold;
        pp_force(k, v[c], v[c + 1], v[c + QX], v[c - 1], v[c - QX], v[c + QX + 1], v[c + QX - 1], v[c - QX - 1],
                 v[c - QX + 1], Fx, Fy);
    } else {
        const uint32_t n = (uint32_t)(wrap(j0 + fy, ny, small) * nx + wrap(i0 + fx, nx, small));
        const uint32_t o = DIR == 0 ? 0u : in.pb;
        Fx = ldf(in.sh, n, SH_FFX * in.pb + o);
        Fy = ldf(in.sh, n, SH_FFY * in.pb + o);
    }
    float sx, sy;
    moment1<L>(fb, sx, sy);
    const float r = rcp(rho);
    hU = fmaf(k.ms, sx, k.half_hdt * Fx) * r;
    hV = fmaf(k.ms, sy, k.half_hdt * Fy) * r;
    hrho = rho;
    unroll<G::NM>([&](auto Z) { hv[decltype(Z)::value] = fb[mq<L, DIR>(decltype(Z)::value)]; });
}

template <class L, class G, int DIR>
__device__ __forceinline__ void s3_face(const KPP &k, const Rs &out, const float *s_v, float *s_fl, int fx, int fy,
                                        int kf, const float *hv, float rho, float U, float V, int i0, int j0, int nx,
                                        int ny)
{
    constexpr int FN = G::FN, NM = G::NM, QX = G::QX;
    const int c = qidx<G>(fx, fy);
    const float vc = s_v[c];
    Site s;
    pp_force(k, vc, s_v[c + 1], s_v[c + QX], s_v[c - 1], s_v[c - QX], s_v[c + QX + 1], s_v[c + QX - 1],
             s_v[c - QX - 1], s_v[c - QX + 1], s.Fx, s.Fy);
    // the tile's own faces (not the last column / row, which the next tile owns) are new state
    if ((DIR == 0 ? fx < G::TX : fy < G::TY) && i0 + fx < nx && j0 + fy < ny)
        stf(out.f, (uint32_t)((j0 + fy) * nx + (i0 + fx)), vc, (L::Q + 2 + DIR) * out.pb);
    s.rho = k.bh * rho;   // feq = bh f_eq
    s.U = U;
    s.V = V;
    float feq[L::Q], fso[L::Q];
    feq_src<L, false>(k, s, feq, fso, [](int p) { return en<L, DIR>(L::first(p)) != 0; });
    float mass = 0.f;
    unroll<NM>([&](auto Z) {
        constexpr int z = decltype(Z)::value, q = mq<L, DIR>(z);
        const float ff = fmaf(k.ch, fso[q], fmaf(k.omb_h, hv[z], feq[q]));
        const float fl = ff * (k.flux * (float)en<L, DIR>(q));
        s_fl[z * FN + kf] = fl;
        mass += fl;
    });
    s_fl[NM * FN + kf] = mass;
}

template <class L, int TX, int TY, int FLUX, bool FORCE, bool TAUU>
__global__ void __launch_bounds__(TX * TY > 256 ? 512 : 256)
    k_pp32_tile(PP32Planes P, PP32Set out_set, KPP k, int ntx, int ntiles)
{
    using G = PTile<L, TX, TY, FLUX == MPD_FLUX_UW ? 2 : 1>;
    constexpr int Q = L::Q, NP = L::NP, NT = G::NT, NIN = G::NIN, RN = G::RN, RX = G::RX, RO = G::RO;
    constexpr int QN = G::QN, QX = G::QX, PO = G::PO, NM = G::NM, FN = G::FN, NXF = G::NXF;
    constexpr int NXA = G::NXA, NYA = G::NYA, EXF = G::EXF, EYF = G::EYF, YC0 = G::YC0;
    __shared__ __align__(16) float s_u[G::NU];
    __shared__ float s_v[2 * QN];                 // this step's face Psi, x then y (Psi-region layout)
    __shared__ float s_vold[FORCE ? 2 * QN : 1];  // the previous step's
    float2 *const s_pair = reinterpret_cast<float2 *>(s_u);
    float *const s_rest = s_u + 2 * NP * RN;
    float *const s_psi = s_rest + (L::REST ? RN : 0);
    float *const s_rho = s_psi + QN;
    float *const s_fl = s_u;  // after the faces have read the bar-plus planes

    const int tile = tile_of_block(blockIdx.x, ntiles);
    if (tile < 0) return;
    const int nx = P.nx, ny = P.ny;
    const int64_t N = P.N;
    const int i0 = (tile % ntx) * TX, j0 = (tile / ntx) * TY;
    const int tid = threadIdx.x;
    const bool small = nx < G::QX || ny < G::QY;
    const Rs in = make_rs(P.f, P.rho, P.tau, N, Q);
    const Rs out = make_rs(out_set.f, out_set.rho, P.tau, N, Q);

    // ---- loads: S0's (the densities and previous face Psi of this thread's
    // Psi-region cells) first, then the S1 inputs of its tile cell.  Loads
    // return in order, so S0's EOS waits only for its own loads while the
    // population loads are in flight; every thread issues the same loads (a
    // thread beyond the tile, or without a second ring item, loads a cell it
    // does not use), so the waits count exactly.
    constexpr int NR0 = QN - NIN, NRI = (NR0 + NT - 1) / NT;  // ring items per thread
    const int ci = tid % TX, cj = tid / TX;
    const bool cell = tid < NIN;
    const bool inside = cell && i0 + ci < nx && j0 + cj < ny;
    const uint32_t cn = (uint32_t)(wrap(j0 + cj, ny, small) * nx + wrap(i0 + ci, nx, small));
    const double crho = ldd(in.rho, cn);
    float cvx = 0.f, cvy = 0.f;
    if (FORCE) {
        cvx = ldf(in.f, cn, (Q + 2) * in.pb);
        cvy = ldf(in.f, cn, (Q + 3) * in.pb);
    }
    // ring items: the threads beyond the tile's cells first, then the others
    double rrho[NRI];
    float rvx[NRI], rvy[NRI];
    int rcx[NRI], rcy[NRI];
    bool rok[NRI];
#pragma unroll
    for (int u = 0; u < NRI; ++u) {
        const int r = tid - NIN + (tid < NIN ? NT : 0) + u * NT;
        rok[u] = r < NR0;
        ring_pos<TX, TY, PO>(rok[u] ? r : 0, rcx[u], rcy[u]);
        const int cx = rcx[u], cy = rcy[u];
        const uint32_t n = (uint32_t)(wrap(j0 + cy, ny, small) * nx + wrap(i0 + cx, nx, small));
        rrho[u] = ldd(in.rho, n);
        if (FORCE) {
            // a cell outside a face-Psi region loads the tile cell's value instead
            // (a line this thread reads anyway), not a line nobody needs
            const bool xin = cx >= -1 && cx <= TX + 1 && cy >= -1 && cy <= TY;
            const bool yin = cx >= -1 && cx <= TX && cy >= -1 && cy <= TY + 1;
            rvx[u] = ldf(in.f, xin ? n : cn, (Q + 2) * in.pb);
            rvy[u] = ldf(in.f, yin ? n : cn, (Q + 3) * in.pb);
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    float cf[Q], cU, cV, ctau = k.tau_u, cFx = 0.f, cFy = 0.f;
#pragma unroll
    for (int q = 0; q < Q; ++q) cf[q] = ldf(in.f, cn, q * in.pb);
    cU = ldf(in.f, cn, Q * in.pb);
    cV = ldf(in.f, cn, (Q + 1) * in.pb);
    if (!TAUU) ctau = ldf(in.sh, cn, SH_TAU * in.pb);
    if (!FORCE) {
        cFx = ldf(in.sh, cn, SH_FX * in.pb);
        cFy = ldf(in.sh, cn, SH_FY * in.pb);
    }
    __builtin_amdgcn_sched_barrier(0);

    // ---- S0: density and signed Psi of the Psi region; previous face Psi ----
    if (cell) {
        const int qi = qidx<G>(ci, cj);
        s_rho[qi] = (float)crho;
        if (FORCE) {
            s_psi[qi] = pp_vsig(k, crho);
            s_vold[qi] = cvx;
            s_vold[QN + qi] = cvy;
        }
    }
#pragma unroll
    for (int u = 0; u < NRI; ++u) {
        if (!rok[u]) continue;
        const int cx = rcx[u], cy = rcy[u], qi = qidx<G>(cx, cy);
        s_rho[qi] = (float)rrho[u];
        if (FORCE) {
            s_psi[qi] = pp_vsig(k, rrho[u]);
            // x-faces fx -1..TX+1, fy -1..TY; y-faces fx -1..TX, fy -1..TY+1
            if (cx >= -1 && cx <= TX + 1 && cy >= -1 && cy <= TY) s_vold[qi] = rvx[u];
            if (cx >= -1 && cx <= TX && cy >= -1 && cy <= TY + 1) s_vold[QN + qi] = rvy[u];
        }
    }
    __syncthreads();

    // ---- S1: cell force and bar-plus (Update_phi_BP, DUGKS2DSolver.cpp:379-401;
    // factors Cell_2D.cpp:120-125, written (1 - bBP) f~ + bBP f_eq + cBP f_So) ----
    auto barplus = [&](int cx, int cy, const float *f, float U, float V, float tau, float &Fx, float &Fy, float *base) {
        const int qi = qidx<G>(cx, cy);
        if (FORCE)
            pp_force(k, s_psi[qi], s_psi[qi + 1], s_psi[qi + QX], s_psi[qi - 1], s_psi[qi - QX], s_psi[qi + QX + 1],
                     s_psi[qi + QX - 1], s_psi[qi - QX - 1], s_psi[qi - QX + 1], Fx, Fy);
        float omb, bbp, cbp;
        if (TAUU) {
            omb = k.omb_u;
            bbp = k.bbp_u;
            cbp = k.cbp_u;
        } else {
            split1(fmaf(-k.bdt, rcp(fmaf(2.f, tau, k.dt)), 1.f), omb, bbp);
            cbp = tau * bbp;
        }
        Site s;
        s.rho = bbp * s_rho[qi];   // feq = bBP f_eq
        s.U = U;
        s.V = V;
        s.Fx = Fx;
        s.Fy = Fy;
        float feq[Q], fso[Q];
        feq_src<L, L::REST>(k, s, feq, fso, [](int) { return true; });
        const int ri = (cy + RO) * RX + (cx + RO);
        float b[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) b[q] = fmaf(cbp, fso[q], fmaf(omb, f[q], feq[q]));
        if constexpr (L::REST) s_rest[ri] = b[0];
#pragma unroll
        for (int p = 0; p < NP; ++p) s_pair[p * RN + ri] = make_float2(b[L::first(p)], b[L::opp(L::first(p))]);
        if (base) {
            const float third = 1.f / 3.f;
#pragma unroll
            for (int q = 0; q < Q; ++q) base[q] = fmaf(third, b[q] - f[q], b[q]);
        }
    };
    float base[Q];
    if (cell) barplus(ci, cj, cf, cU, cV, ctau, cFx, cFy, base);
    for (int r = tid - NIN + (tid < NIN ? NT : 0); r < RN - NIN; r += NT) {
        int cx, cy;
        ring_pos<TX, TY, RO>(r, cx, cy);
        const uint32_t n = (uint32_t)(wrap(j0 + cy, ny, small) * nx + wrap(i0 + cx, nx, small));
        float f[Q], Fx = 0.f, Fy = 0.f, tau = k.tau_u;
#pragma unroll
        for (int q = 0; q < Q; ++q) f[q] = ldf(in.f, n, q * in.pb);
        const float U = ldf(in.f, n, Q * in.pb), V = ldf(in.f, n, (Q + 1) * in.pb);
        if (!TAUU) tau = ldf(in.sh, n, SH_TAU * in.pb);
        if (!FORCE) {
            Fx = ldf(in.sh, n, SH_FX * in.pb);
            Fy = ldf(in.sh, n, SH_FY * in.pb);
        }
        barplus(cx, cy, f, U, V, tau, Fx, Fy, nullptr);
    }
    __syncthreads();

    // ---- S2: faces ----
    // pass A: tile x-face tid (fx = tid % TX, fy = tid / TX; then the last column);
    // pass B: tile y-face tid; pass C: the remaining tile x-faces and the Psi-only
    // x-faces from thread 0, the remaining tile y-faces and the Psi-only y-faces
    // from the first thread of a later wave (each wave one direction)
    float av[NM], bv[NM], cv[NM];
    float arho = 1.f, aU = 0.f, aV = 0.f, brho = 1.f, bU = 0.f, bV = 0.f, crh = 1.f, cU2 = 0.f, cV2 = 0.f;
    const bool pa = tid < NXA, pb = tid < NYA;
    int afx = 0, afy = 0, bfx = 0, bfy = 0;
    if (pa) {
        afx = tid < NIN ? tid % TX : TX;
        afy = tid < NIN ? tid / TX : tid - NIN;
        s2_face<L, G, FLUX, FORCE, 0>(k, in, s_pair, s_rest, s_vold, s_v, afx, afy, true, i0, j0, nx, ny, small, av,
                                      arho, aU, aV);
    }
    __builtin_amdgcn_sched_barrier(0);
    if (pb) {
        bfx = tid % TX;
        bfy = tid / TX;
        s2_face<L, G, FLUX, FORCE, 1>(k, in, s_pair, s_rest, s_vold + QN, s_v + QN, bfx, bfy, true, i0, j0, nx, ny,
                                      small, bv, brho, bU, bV);
    }
    int cfx = 0, cfy = 0;
    const bool cx_ = tid < EXF + G::NXR, cy_ = tid >= YC0 && tid < G::NC;
    const bool ctile = cx_ ? tid < EXF : (cy_ && tid - YC0 < EYF);
    if (cx_) {
        if (tid < EXF) {
            const int t = NXA + tid;  // x-faces beyond the first pass: the last column
            cfx = t < NIN ? t % TX : TX;
            cfy = t < NIN ? t / TX : t - NIN;
        } else {
            ring_face<G, 0>(tid - EXF, cfx, cfy);
        }
        s2_face<L, G, FLUX, FORCE, 0>(k, in, s_pair, s_rest, s_vold, s_v, cfx, cfy, ctile, i0, j0, nx, ny, small, cv,
                                      crh, cU2, cV2);
    } else if (cy_) {
        const int t = tid - YC0;
        if (t < EYF) {
            const int u = NYA + t;  // y-faces beyond the first pass: the last row
            cfx = u % TX;
            cfy = u / TX;
        } else {
            ring_face<G, 1>(t - EYF, cfx, cfy);
        }
        s2_face<L, G, FLUX, FORCE, 1>(k, in, s_pair, s_rest, s_vold + QN, s_v + QN, cfx, cfy, ctile, i0, j0, nx, ny,
                                      small, cv, crh, cU2, cV2);
    }
    __syncthreads();

    // ---- S3: face force, BGK and flux of the tile faces ----
    // flux slot of a tile face: x-face (fx, fy) at fx < TX ? fy TX + fx : TX TY + fy, y-faces after them
    auto xslot = [&](int fx, int fy) { return fx < TX ? fy * TX + fx : NIN + fy; };
    if (pa) s3_face<L, G, 0>(k, out, s_v, s_fl, afx, afy, xslot(afx, afy), av, arho, aU, aV, i0, j0, nx, ny);
    __builtin_amdgcn_sched_barrier(0);
    if (pb) s3_face<L, G, 1>(k, out, s_v + QN, s_fl, bfx, bfy, NXF + bfy * TX + bfx, bv, brho, bU, bV, i0, j0, nx, ny);
    if (ctile) {
        if (cx_) s3_face<L, G, 0>(k, out, s_v, s_fl, cfx, cfy, xslot(cfx, cfy), cv, crh, cU2, cV2, i0, j0, nx, ny);
        else s3_face<L, G, 1>(k, out, s_v + QN, s_fl, cfx, cfy, NXF + cfy * TX + cfx, cv, crh, cU2, cV2, i0, j0, nx, ny);
    }
    __syncthreads();

    // ---- S4: f~' = b + (b - f~)/3 + dt/V (W - E + S - N), FP64 density ----
    if (!inside) return;
    const int kW = cj * TX + ci, kE = ci + 1 < TX ? kW + 1 : NIN + cj;
    const int kS = NXF + cj * TX + ci, kN = kS + TX;
    float ft[Q];
    unroll<Q>([&](auto QQ) {
        constexpr int q = decltype(QQ)::value, zx = slot<L, 0>(q), zy = slot<L, 1>(q);
        float s = 0.f;
        if constexpr (zx >= 0) s += s_fl[zx * FN + kW] - s_fl[zx * FN + kE];
        if constexpr (zy >= 0) s += s_fl[zy * FN + kS] - s_fl[zy * FN + kN];
        ft[q] = fmaf(k.dtv, s, base[q]);
    });
    const double dm = ((double)s_fl[NM * FN + kW] - (double)s_fl[NM * FN + kE]) +
                      ((double)s_fl[NM * FN + kS] - (double)s_fl[NM * FN + kN]);
    const double rho_d = crho + (double)k.dtv * dm;
    if constexpr (L::REST) {
        double rest = 0.0;
#pragma unroll
        for (int q = 1; q < Q; ++q) rest += (double)ft[q];
        ft[0] = (float)(rho_d - rest);
    } else {
        double sum = 0.0;
#pragma unroll
        for (int q = 0; q < Q; ++q) sum += (double)ft[q];
        const double d4 = 0.25 * (rho_d - sum);
#pragma unroll
        for (int q = 4; q < 8; ++q) ft[q] = (float)((double)ft[q] + d4);
    }
    const uint32_t n = (uint32_t)((j0 + cj) * nx + (i0 + ci));
#pragma unroll
    for (int q = 0; q < Q; ++q) stf(out.f, n, ft[q], q * out.pb);
    std_(out.rho, n, rho_d);
    // Update_MacroVar (D2Q9PP.cpp:89-109): u = (ms sum e f~ + hdt F) / rho
    float sx, sy;
    moment1<L>(ft, sx, sy);
    const float r = rcp((float)rho_d);
    stf(out.f, n, fmaf(k.ms, sx, k.hdt * cFx) * r, Q * out.pb);
    stf(out.f, n, fmaf(k.ms, sy, k.hdt * cFy) * r, (Q + 1) * out.pb);
}

// ---------------- one launch per phase (D2Q9) ----------------
// The phases above as four kernels over global planes (scratch of PP32Planes),
// on the same state: the paths can be switched between steps.

inline unsigned blocks(size_t n) { return (unsigned)((n + 255) / 256); }
__device__ __forceinline__ int wrapp(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// cell Psi of the density (FP64 EOS), signed by scG
__global__ void __launch_bounds__(256) k_pp32_psi(PP32Planes P, KPP k)
{
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (size_t)P.N) return;
    P.psi[n] = pp_vsig(k, P.rho[n]);
}

// the force of signed-Psi plane v at site (i, j) of an nx x ny lattice
__device__ __forceinline__ void plane_force(const KPP &k, const float *v, int nx, int ny, int i, int j, float &Fx,
                                            float &Fy)
{
    auto S = [&](int a, int b) { return v[(size_t)wrapp(b, ny) * nx + wrapp(a, nx)]; };
    pp_force(k, S(i, j), S(i + 1, j), S(i, j + 1), S(i - 1, j), S(i, j - 1), S(i + 1, j + 1), S(i - 1, j + 1),
             S(i - 1, j - 1), S(i + 1, j - 1), Fx, Fy);
}

// phase 1: cell force (FORCE: from Psi, kept in Fx/Fy; else the stored one)
// and bar-plus of every cell
template <bool FORCE>
__global__ void __launch_bounds__(256) k_pp32_phase1(PP32Planes P, KPP k)
{
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t N = P.N;
    if (n >= N) return;
    float Fx, Fy;
    if (FORCE) {
        plane_force(k, P.psi, P.nx, P.ny, (int)(n % P.nx), (int)(n / P.nx), Fx, Fy);
        P.Fx[n] = Fx;
        P.Fy[n] = Fy;
    } else {
        Fx = P.Fx[n];
        Fy = P.Fy[n];
    }
    const float tau = P.tau[n];
    float omb, bbp;
    split1(fmaf(-k.bdt, rcp(fmaf(2.f, tau, k.dt)), 1.f), omb, bbp);
    const float cbp = tau * bbp;
    Site s;
    s.rho = bbp * (float)P.rho[n];   // feq = bBP f_eq
    s.U = pp32_U(P.f, 9, N)[n];
    s.V = pp32_V(P.f, 9, N)[n];
    s.Fx = Fx;
    s.Fy = Fy;
    float feq[9], fso[9];
    feq_src<L9, true>(k, s, feq, fso, [](int) { return true; });
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const float f = P.f[q * N + n];
        P.fbp[q * N + n] = fmaf(cbp, fso[q], fmaf(omb, f, feq[q]));
    }
}

// faces, pass 1: face values (kept in the flux planes), face macros with the
// face's previous force, and face Psi
template <bool FORCE>
__global__ void __launch_bounds__(256) k_pp32_face_psi(PP32Planes P, KPP k)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t N = P.N;
    if (t >= 2 * N) return;
    const int dir = t < N ? 0 : 1;
    const size_t fi = dir ? t - N : t;
    const int nx = P.nx, ny = P.ny;
    const int i = (int)(fi % nx), j = (int)(fi / nx);
    auto B = [&](int q, int a, int b) { return P.fbp[(size_t)q * N + (size_t)wrapp(b, ny) * nx + wrapp(a, nx)]; };
    float fb[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const int ox = dir == 0 ? i - 1 : i, oy = dir == 0 ? j : j - 1;
        const int e_n = dir == 0 ? ex9(q) : ey9(q), e_t = dir == 0 ? ey9(q) : ex9(q);
        const float bo = B(q, ox, oy), bn = B(q, i, j);
        float v = e_n > 0 ? fmaf(k.rb[0], bn, k.ra[0] * bo) : (e_n < 0 ? fmaf(k.ra[0], bn, k.rb[0] * bo) : 0.5f * (bo + bn));
        if (e_t != 0) {
            const int tx = dir == 0 ? 0 : 1, ty = dir == 0 ? 1 : 0;
            const float tr = (B(q, i + tx, j + ty) - B(q, i - tx, j - ty)) + (B(q, ox + tx, oy + ty) - B(q, ox - tx, oy - ty));
            v = fmaf(e_t > 0 ? -k.rt4 : k.rt4, tr, v);
        }
        fb[q] = v;
        P.ffl[(size_t)q * 2 * N + t] = v;  // the flux pass reads it back and overwrites it with the flux
    }
    float Fx, Fy;
    if (FORCE) plane_force(k, pp32_fv(P.f, 9, N) + (dir ? N : 0), nx, ny, i, j, Fx, Fy);
    else {
        Fx = P.ffx[t];
        Fy = P.ffy[t];
    }
    const float rho = face_rho<L9>(fb);
    float sx, sy;
    moment1<L9>(fb, sx, sy);
    const float r = rcp(rho);
    P.frho[t] = rho;
    P.fU[t] = fmaf(k.ms, sx, k.half_hdt * Fx) * r;
    P.fV[t] = fmaf(k.ms, sy, k.half_hdt * Fy) * r;
    P.fpsi[t] = pp_vsig(k, (double)rho);
}

// faces, pass 2: face force from the face-lattice gradient of Psi, BGK half
// step and micro-flux; the face Psi becomes the set's face state
__global__ void __launch_bounds__(256) k_pp32_face_flux(PP32Planes P, KPP k)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t N = P.N;
    if (t >= 2 * N) return;
    const int dir = t < N ? 0 : 1;
    const size_t fi = dir ? t - N : t;
    const int nx = P.nx, ny = P.ny;
    Site s;
    plane_force(k, P.fpsi + (dir ? N : 0), nx, ny, (int)(fi % nx), (int)(fi / nx), s.Fx, s.Fy);
    pp32_fv(P.f, 9, N)[t] = P.fpsi[t];
    s.rho = k.bh * P.frho[t];   // feq = bh f_eq
    s.U = P.fU[t];
    s.V = P.fV[t];
    float feq[9], fso[9];
    if (dir == 0) feq_src<L9, false>(k, s, feq, fso, [](int p) { return ex9(L9::first(p)) != 0; });
    else feq_src<L9, false>(k, s, feq, fso, [](int p) { return ey9(L9::first(p)) != 0; });
    float mass = 0.f;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const int e_n = dir == 0 ? ex9(q) : ey9(q);
        float fl = 0.f;
        if (e_n != 0) {
            const float v = P.ffl[(size_t)q * 2 * N + t];
            const float ff = fmaf(k.ch, fso[q], fmaf(k.omb_h, v, feq[q]));
            fl = ff * (k.flux * (float)e_n);
        }
        P.ffl[(size_t)q * 2 * N + t] = fl;
        mass += fl;
    }
    P.fmass[t] = mass;
}

// Update_phi_T and Update_MacroVar with the FP64 density
__global__ void __launch_bounds__(256) k_pp32_update(PP32Planes P, KPP k)
{
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t N = P.N;
    if (n >= N) return;
    const int nx = P.nx, ny = P.ny;
    const int i = (int)(n % nx), j = (int)(n / nx);
    const size_t kE = (size_t)j * nx + wrapp(i + 1, nx), kW = n;
    const size_t kN = N + (size_t)wrapp(j + 1, ny) * nx + i, kS = N + n;
    float ft[9];
    const float third = 1.f / 3.f;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const float *fl = P.ffl + (size_t)q * 2 * N;
        const float s = (fl[kW] - fl[kE]) + (fl[kS] - fl[kN]);
        const float b = P.fbp[q * N + n], f = P.f[q * N + n];
        ft[q] = fmaf(k.dtv, s, fmaf(third, b - f, b));
    }
    const double dm = ((double)P.fmass[kW] - (double)P.fmass[kE]) + ((double)P.fmass[kS] - (double)P.fmass[kN]);
    const double rho_d = P.rho[n] + (double)k.dtv * dm;
    double rest = 0.0;
#pragma unroll
    for (int q = 1; q < 9; ++q) rest += (double)ft[q];
    ft[0] = (float)(rho_d - rest);
#pragma unroll
    for (int q = 0; q < 9; ++q) P.f[q * N + n] = ft[q];
    P.rho[n] = rho_d;
    float sx, sy;
    moment1<L9>(ft, sx, sy);
    const float r = rcp((float)rho_d);
    pp32_U(P.f, 9, N)[n] = fmaf(k.ms, sx, k.hdt * P.Fx[n]) * r;
    pp32_V(P.f, 9, N)[n] = fmaf(k.ms, sy, k.hdt * P.Fy[n]) * r;
}

// ---------------- conversions ----------------

// explicit FP64 state -> FP32 set and the shared planes (its forces)
__global__ void __launch_bounds__(256) k_pp32_from64(PP32Planes P, DevState S)
{
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t N = P.N;
    if (n >= N) return;
    const int q = P.q;
    const double *m = S.m;
    for (int j = 0; j < q; ++j) P.f[j * N + n] = (float)S.f[j * N + n];
    P.rho[n] = m[MPD_M_RHO * N + n];
    pp32_U(P.f, q, N)[n] = (float)m[MPD_M_U * N + n];
    pp32_V(P.f, q, N)[n] = (float)m[MPD_M_V * N + n];
    P.tau[n] = (float)m[MPD_M_TAU * N + n];
    P.Fx[n] = (float)m[MPD_M_FX * N + n];
    P.Fy[n] = (float)m[MPD_M_FY * N + n];
    for (size_t t = n; t < 2 * N; t += N) {
        P.ffx[t] = (float)S.ffx[t];
        P.ffy[t] = (float)S.ffy[t];
        pp32_fv(P.f, q, N)[t] = 0.f;
    }
}

// the forces of a stepped state (the reference's last phase, Update_PseudoForce
// of the cells, and the face forces the last step computed), into the shared planes
__global__ void __launch_bounds__(256) k_pp32_forces(PP32Planes P, KPP k)
{
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t N = P.N;
    if (n >= N) return;
    const int nx = P.nx, ny = P.ny, i = (int)(n % nx), j = (int)(n / nx);
    auto V = [&](int a, int b) { return pp_vsig(k, P.rho[(size_t)wrapp(b, ny) * nx + wrapp(a, nx)]); };
    float Fx, Fy;
    pp_force(k, V(i, j), V(i + 1, j), V(i, j + 1), V(i - 1, j), V(i, j - 1), V(i + 1, j + 1), V(i - 1, j + 1),
             V(i - 1, j - 1), V(i + 1, j - 1), Fx, Fy);
    P.Fx[n] = Fx;
    P.Fy[n] = Fy;
    const float *fv = pp32_fv(P.f, P.q, N);
    for (int d = 0; d < 2; ++d) {
        plane_force(k, fv + d * N, nx, ny, i, j, Fx, Fy);
        P.ffx[d * N + n] = Fx;
        P.ffy[d * N + n] = Fy;
    }
}

// FP32 set -> explicit FP64 state (p, Psi, scG from the FP64 EOS of the density)
__global__ void __launch_bounds__(256) k_pp32_to64(PP32Planes P, DevState S, Consts C, KPP k)
{
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t N = P.N;
    if (n >= N) return;
    const int q = P.q;
    double *m = S.m;
    for (int j = 0; j < q; ++j) S.f[j * N + n] = (double)P.f[j * N + n];
    const double rho = P.rho[n];
    m[MPD_M_RHO * N + n] = rho;
    m[MPD_M_U * N + n] = (double)pp32_U(P.f, q, N)[n];
    m[MPD_M_V * N + n] = (double)pp32_V(P.f, q, N)[n];
    m[MPD_M_P * N + n] = C.pp_k * carnahan_starling(rho, C.t_cs);
    m[MPD_M_FX * N + n] = (double)P.Fx[n];
    m[MPD_M_FY * N + n] = (double)P.Fy[n];
    const float v = pp_vsig(k, rho);
    m[MPD_M_PSI * N + n] = (double)fabsf(v);
    m[MPD_M_SCG * N + n] = v > 0.f ? 1.0 : (v < 0.f ? -1.0 : 0.0);
    m[MPD_M_TAU * N + n] = (double)P.tau[n];
    for (size_t t = n; t < 2 * N; t += N) {
        S.ffx[t] = (double)P.ffx[t];
        S.ffy[t] = (double)P.ffy[t];
    }
}

}  // namespace

int pp32_set_planes(int q) { return q + 4; }

int pp32_alloc(PP32State &St, void **mem, int64_t *bytes, const Consts &C, int64_t N, int q)
{
    // two sets of: FP64 density [N]; float block (q + 4) [N]; then the shared
    // planes tau, Fx, Fy [N], ffx, ffy [2N]; D2Q9 also the per-phase scratch:
    // bar-plus 9 [N], cell scG Psi [N], face scG Psi, rho, U, V, mass 5 [2N], face values 9 [2N]
    const size_t n = (size_t)N;
    const size_t set_f = (size_t)pp32_set_planes(q) * n;
    const size_t scratch_f = q == 9 ? 10 * n + 10 * n + 18 * n : 0;
    const size_t bytes_total = 2 * n * sizeof(double) + (2 * set_f + SH_PLANES * n + scratch_f) * sizeof(float);
    void *base = nullptr;
    if (hipMalloc(&base, bytes_total) != hipSuccess) return -1;
    *mem = base;
    *bytes = (int64_t)bytes_total;
    PP32Planes &P = St.P;
    P = PP32Planes{};
    P.nx = C.nx;
    P.ny = C.ny;
    P.q = q;
    P.N = N;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
        St.cus = cus;
    double *d = (double *)base;
    float *p = (float *)(d + 2 * n);
    P.rho = d;
    P.f = p;
    St.other.rho = d + n;
    St.other.f = p + set_f;
    p += 2 * set_f;
    // shared planes, consecutive (SH_* offsets)
    P.tau = p;
    P.Fx = p + n;
    P.Fy = p + 2 * n;
    P.ffx = p + 3 * n;
    P.ffy = p + 5 * n;
    p += SH_PLANES * n;
    if (q == 9) {
        P.fbp = p; p += 9 * n;
        P.psi = p; p += n;
        P.fpsi = p; p += 2 * n;
        P.frho = p; p += 2 * n;
        P.fU = p; p += 2 * n;
        P.fV = p; p += 2 * n;
        P.fmass = p; p += 2 * n;
        P.ffl = p;
    }
    return 0;
}

void pp32_from64(const PP32State &St, const DevState &S, hipStream_t st)
{
    hipLaunchKernelGGL(k_pp32_from64, dim3(blocks(St.P.N)), dim3(256), 0, st, St.P, S);
}

void pp32_to64(const PP32State &St, const DevState &S, const Consts &C, bool stepped, hipStream_t st)
{
    const KPP k = make_kpp(C, St.P.q, 0.0);
    if (stepped) hipLaunchKernelGGL(k_pp32_forces, dim3(blocks(St.P.N)), dim3(256), 0, st, St.P, k);
    hipLaunchKernelGGL(k_pp32_to64, dim3(blocks(St.P.N)), dim3(256), 0, st, St.P, S, C, k);
}

int pp32_tile_auto(int nx, int ny, int q, int cus)
{
    // D2V16: 16x15 tiles.  D2Q9: 32x16 tiles (per cell about 10% less halo work
    // than 16x15) unless fewer than half the CUs would get one.  Measured on
    // MI355X, us per step 16x15 / 32x16: 128^2 9.6 / 10.4, 256^2 8.9 / 8.6,
    // 512^2 15.2 / 11.5, 1024^2 41 / 35, 2048^2 172 / 149.
    if (q != 9) return MPD_TILE_16X15;
    return 2 * (int64_t)((nx + 31) / 32) * ((ny + 15) / 16) >= (cus > 0 ? cus : 256) ? MPD_TILE_32X16
                                                                                        : MPD_TILE_16X15;
}

template <class L, int TX, int TY, int FLUX>
void pp32_tile_launch(PP32State &St, const KPP &k, bool stepped, hipStream_t st)
{
    const PP32Planes &P = St.P;
    const int ntx = (P.nx + TX - 1) / TX, nty = (P.ny + TY - 1) / TY, ntiles = ntx * nty;
    const dim3 g((unsigned)(((ntiles + 7) / 8) * 8)), b(PTile<L, TX, TY, 1>::NT);
    const bool u = k.tau_u > 0.f;
    if (stepped && u) hipLaunchKernelGGL((k_pp32_tile<L, TX, TY, FLUX, true, true>), g, b, 0, st, P, St.other, k, ntx, ntiles);
    else if (stepped) hipLaunchKernelGGL((k_pp32_tile<L, TX, TY, FLUX, true, false>), g, b, 0, st, P, St.other, k, ntx, ntiles);
    else if (u) hipLaunchKernelGGL((k_pp32_tile<L, TX, TY, FLUX, false, true>), g, b, 0, st, P, St.other, k, ntx, ntiles);
    else hipLaunchKernelGGL((k_pp32_tile<L, TX, TY, FLUX, false, false>), g, b, 0, st, P, St.other, k, ntx, ntiles);
}

void pp32_step(PP32State &St, const Consts &C, bool stepped, hipStream_t st)
{
    PP32Planes &P = St.P;
    const size_t N = P.N;
    const KPP k = make_kpp(C, P.q, St.tau_u);
    if (St.phases && P.q == 9) {
        if (stepped) {
            hipLaunchKernelGGL(k_pp32_psi, dim3(blocks(N)), dim3(256), 0, st, P, k);
            hipLaunchKernelGGL(k_pp32_phase1<true>, dim3(blocks(N)), dim3(256), 0, st, P, k);
            hipLaunchKernelGGL(k_pp32_face_psi<true>, dim3(blocks(2 * N)), dim3(256), 0, st, P, k);
        } else {
            hipLaunchKernelGGL(k_pp32_phase1<false>, dim3(blocks(N)), dim3(256), 0, st, P, k);
            hipLaunchKernelGGL(k_pp32_face_psi<false>, dim3(blocks(2 * N)), dim3(256), 0, st, P, k);
        }
        hipLaunchKernelGGL(k_pp32_face_flux, dim3(blocks(2 * N)), dim3(256), 0, st, P, k);
        hipLaunchKernelGGL(k_pp32_update, dim3(blocks(N)), dim3(256), 0, st, P, k);
        return;
    }
    const int tile = St.tile ? St.tile : pp32_tile_auto(P.nx, P.ny, P.q, St.cus);
    if (P.q == 16) pp32_tile_launch<L16, 16, 15, MPD_FLUX_CD>(St, k, stepped, st);
    else if (C.flux == MPD_FLUX_UW) pp32_tile_launch<L9, 16, 15, MPD_FLUX_UW>(St, k, stepped, st);  // 16x15 only
    else if (tile == MPD_TILE_32X16) pp32_tile_launch<L9, 32, 16, MPD_FLUX_CD>(St, k, stepped, st);
    else pp32_tile_launch<L9, 16, 15, MPD_FLUX_CD>(St, k, stepped, st);
    // the written set becomes the current one
    const PP32Set cur{P.rho, P.f};
    P.rho = St.other.rho;
    P.f = St.other.f;
    St.other = cur;
}

}  // namespace mpd
