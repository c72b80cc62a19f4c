/*
 * CPU FP64 restatement of the reference's 2D DUGKS time step - THE ORACLE.
 *
 * Test infrastructure only.  Nothing in the product links, loads or calls
 * this file; it is the reference implementation used by tests/, by __graft_entry__.smoke() and
 * by bench.py's CPU-baseline leg.
 *
 * It restates `DUGKS2DSolver()`'s loop body (Ark/Cpp/DUGKS2DSolver.cpp:164-328)
 * on a structured nx*ny grid, keeping the reference's phase order and its
 * floating-point operation order so that it tracks the reference to round-off:
 *
 *   phase 1  Update_phi_Eq / Update_DVDF_Source / Update_phi_BP      (:175-183)
 *   phase 2  WallShadowC_fBP (walls only)                            (:185-189)
 *   phase 3  Grad_VS_6points (UW only)                               (:197-206)
 *   phase 4  Flux_2D on every face (+ Wall_3_BB on wall faces)       (:219-250)
 *   phase 6  Update_phiFlux_h                                        (:269-273)
 *   phase 8  Update_phi_T                                            (:280-282)
 *   phase 9  Update_MacroVar                                         (:284-288)
 *   phase 10 MacroSource (Allen-Cahn)                                (:289-295)
 *   phase 11 Update_PseudoForce(cell) (pseudopotential)              (:297-303)
 *
 * Face convention: x-face (i,j) is the west face of cell (i,j) with normal +x,
 * owner (i-1,j), neighbour (i,j); y-face (i,j) is the south face of cell (i,j)
 * with normal +y, owner (i,j-1).  The reference stores each periodic boundary
 * face twice (zones 12 and 8) with opposite normals, each owned by its mesh
 * cell.  With CD faces both copies evaluate to the same bits as the single
 * structured face here, because every operation that differs is an exact sign
 * flip (see face_recon()).  With UW faces the xi.n = 0 average adds the owner
 * first, so the copies round apart: the structured face at index 0 is the
 * zone-12 copy, and the zone-8 copy (make_face at i = nx / j = ny) is evaluated
 * as well for the last column / row (cbhl, cbfl; cbpsi in the reference face
 * order).
 *
 * The pseudopotential face force has two modes.  The reference writes each
 * face's Psi and reads its neighbour faces' Psi inside the same OpenMP loop
 * (Ark/Cpp/DUGKS2DSolver.cpp:219-230, 733-746), so its result depends on the
 * face visiting order.  The default (Jacobi) mode computes all face Psi first
 * and the face forces second, the order the GPU uses.  The reference-order
 * mode (ora_set_face_order) reproduces one thread's visit instead: a stencil
 * neighbour visited earlier contributes its new Psi, one visited later the Psi
 * it kept from the previous step; the two stored copies of each periodic face
 * keep their own forces and fluxes (see pp_forces_ref()).  Both modes share
 * every other function.
 */
#include "dugks_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define Q9 9
#define QMAX 16

/* D2Q9 lattice, Ark/Lib/D2Q9.h:24-28 and weights Ark/Lib/D2Q9HCZ.cpp:15-19 */
static const double E9X[Q9] = {0, 1, 1, 0, -1, -1, -1, 0, 1};
static const double E9Y[Q9] = {0, 0, 1, 1, 1, 0, -1, -1, -1};
static const int BB9[Q9] = {0, 5, 6, 7, 8, 1, 2, 3, 4};
static const double W9[Q9] = {4.0 / 9.0, 1.0 / 9.0, 1.0 / 36.0, 1.0 / 9.0, 1.0 / 36.0,
                              1.0 / 9.0, 1.0 / 36.0, 1.0 / 9.0, 1.0 / 36.0};

struct ora {
    mpd_config c;
    int nx, ny, N, walls;
    /* derived constants (formulas as in Ark/Constant/ZeroNamespace.h) */
    int Q, pp;                /* populations per cell; pseudopotential model (D2Q9 or D2V16) */
    double hdt, lambda0, aphi, bphi, beta, kappa, kp, taum, xi[2][QMAX], dtv;
    double v16a[QMAX][QMAX];  /* D2V16 equilibrium matrix (iArkC) */
    double ah_h, bh_h, ch_h, abp_h, bbp_h, cbp_h; /* h factors: constant tau_h */
    double ah_pp, bh_pp, ch_pp;                    /* PP face factors (never updated) */
    /* state */
    double *h, *f, *m;        /* [Q][N], [Q][N], [MPD_M_COUNT][N] */
    double *ffx, *ffy, *fpsi; /* PP face force / Psi: x-faces [N] then y-faces [N] */
    /* scratch */
    double *hbp, *fbp, *hbpx, *hbpy, *fbpx, *fbpy;
    double *hfl, *ffl;        /* fluxes: x-faces [Q][N] then y-faces [Q][NY1] */
    double *gm;               /* wall ghosts [MPD_M_COUNT][2][nx] */
    double *ghbp, *gfbp;      /* wall ghost bar-plus [Q][2][nx] */
    int nyf;                  /* rows of y-faces: ny (periodic) or ny+1 (walls) */
    /* pseudopotential, reference face order (ora_set_face_order) */
    int ref_order;
    double *fpsi_old;         /* face Psi kept from the previous step [2N] (0 at start:
                                 the initialisers never set it, Ark/Cpp/Preprocess.cpp:337-401) */
    double *cbfx, *cbfy;      /* force of the second copy (zone 8) of each periodic face:
                                 x-faces i = 0 by row [ny], then y-faces j = 0 by column [nx] */
    double *cbfl;             /* flux of those copies [Q][ny + nx] */
    double *cbpsi, *cbpsi_old;/* Psi of those copies, this step and the previous one [ny + nx]
                                 (with UW faces their reconstruction, and so Psi, can
                                 differ from the first copy's in the last bits) */
    double *cbhl;             /* h flux of those copies (Allen-Cahn UW) [Q][ny + nx] */
    int one_copy;             /* test knob (ora_set_one_copy): UW without the second copies */
};

/* ---------------- small helpers ---------------- */

static inline int wrap(int i, int n) { return (i % n + n) % n; }

static inline void set_zero(double *d)
{ /* SetZero, Ark/Header/Mesh_2D.h:26-30 */
    if (*d < 1.0E-14 && *d > -1.0E-14) *d = 0.0;
}

/* value of macro field `fld` at cell (i,j); j = -1 / ny address wall ghosts */
static inline double M(const ora *o, int fld, int i, int j)
{
    i = wrap(i, o->nx);
    if (j < 0 || j >= o->ny) {
        if (!o->walls) j = wrap(j, o->ny);
        else return o->gm[((size_t)fld * 2 + (j < 0 ? 0 : 1)) * o->nx + i];
    }
    return o->m[(size_t)fld * o->N + (size_t)j * o->nx + i];
}

static inline double BP(const ora *o, const double *bp, const double *gbp, int q, int i, int j)
{
    i = wrap(i, o->nx);
    if (j < 0 || j >= o->ny) {
        if (!o->walls) j = wrap(j, o->ny);
        else return gbp[((size_t)q * 2 + (j < 0 ? 0 : 1)) * o->nx + i];
    }
    return bp[(size_t)q * o->N + (size_t)j * o->nx + i];
}

static inline double *MP(ora *o, int fld) { return o->m + (size_t)fld * o->N; }

/* Carnahan-Starling EOS with a = 1, b = 4, T = Tr*0.0943 (Ark/Lib/nonidealEoS.h:59-73) */
static double carnahan_starling(const ora *o, double rho)
{
    double t_cs = o->c.pp_tr * 0.0943;
    return rho * 1.0 * t_cs * (1.0 + rho + rho * rho - rho * rho * rho) /
               ((1.0 - rho) * (1.0 - rho) * (1.0 - rho)) -
           1.0 * rho * rho;
}

/* MacroQuantity::calcPsi, Ark/Header/Macroscopic.h:64-68 */
static void calc_psi(double p, double rho, double rt, double *scg, double *psi)
{
    double s = (double)((p > rho * rt) - (p < rho * rt));
    *scg = s;
    *psi = sqrt(2 * (p - rho * rt) / s);
}

/* D2V16 (Ark/Lib/D2V16.h:14-20, Ark/Lib/D2V16.cpp:11-75): c = MaSpan = 1.2,
 * eta = (3 + nK) c, nK = 0.  Velocities in four shells of four (axis c,
 * diagonal c, axis 2c, diagonal 2c); row j of the equilibrium matrix iArkC from
 * the closed-form coefficients of D2V16.cpp:17-35, written per shell with the
 * direction signs (sx, sy) of velocity j.  e[j] = |xi|^2 (+ eta^2, first shell). */
static void d2v16_tables(double a[QMAX][QMAX], double *xi_x, double *xi_y, double *e)
{
    const double c = 1.2, et = (3.0 + 0) * c;
    const double c2 = c * c, e2 = et * et, c2e2 = c2 * e2;
    const double A = 1.0 / (4.0 * e2), B = (-21.0 * c2 - e2) / (48.0 * c2e2);
    const double D = (7.0 * c2 + e2) / (32.0 * c2e2), E = (-3.0 * c2 - e2) / (96.0 * c2e2);
    const double F = 1.0 / (3.0 * c2);
    const double G = (3.0 * c2 - 5.0 * e2) / (4.0 * e2 * (e2 - 3.0 * c2));
    const double H = 3.0 * (c2 + e2) / (4.0 * e2 * (e2 - 3.0 * c2));
    const double I = (21.0 * c2 - 11.0 * e2) / (48.0 * c2e2);
    const double J = (21.0 * c2 * c2 - 32.0 * c2e2 + 11.0 * e2 * e2) / (32.0 * c2e2 * (e2 - 3.0 * c2));
    const double K = (21.0 * c2 * c2 - 36.0 * c2e2 + 7.0 * e2 * e2) / (32.0 * c2e2 * (e2 - 3.0 * c2));
    const double L = (3.0 * c2 - 5.0 * e2) / (96.0 * c2e2);
    const double Nn = 1.0 / (2.0 * c * e2), P = (3.0 * c2 - e2) / (12.0 * c2e2 * c);
    const double Qq = (c2 + e2) / (16.0 * c2e2 * c), R = (c2 - e2) / (16.0 * c2e2 * c);
    const double S = 1.0 / (24.0 * c2 * c2), T = 1.0 / (4.0 * c2 * (e2 - 3.0 * c2));
    const double U = (c2 - e2) / (32.0 * c2 * c2 * (e2 - 3.0 * c2));
    const double V = (5.0 * c2 - e2) / (32.0 * c2 * c2 * (e2 - 3.0 * c2));
    /* direction signs: axis shells E N W S, diagonal shells NE NW SW SE */
    static const int AX[4] = {1, 0, -1, 0}, AY[4] = {0, 1, 0, -1};
    static const int DX[4] = {1, -1, -1, 1}, DY[4] = {1, 1, -1, -1};
    for (int j = 0; j < 16; ++j) {
        int sh = j / 4, r = j % 4;
        double x = sh % 2 == 0 ? AX[r] : DX[r], y = sh % 2 == 0 ? AY[r] : DY[r];
        double xx = x * x, yy = y * y, sp = sh < 2 ? c : 2 * c;
        double *w = a[j];
        if (xi_x) xi_x[j] = sp * x;
        if (xi_y) xi_y[j] = sp * y;
        if (e) e[j] = sh == 0 ? c2 + e2 : sh == 1 ? 2 * c2 : sh == 2 ? 4 * c2 : 8 * c2;
        for (int m = 0; m < 16; ++m) w[m] = 0.0;
        if (sh == 0) {
            w[3] = A; w[5] = xx * G + yy * H; w[6] = xx * H + yy * G;
            w[7] = x * Nn; w[8] = y * Nn; w[9] = -x * Nn; w[10] = -y * Nn; w[11] = -y * Nn; w[12] = -x * Nn;
            w[14] = (xx - yy) * T; w[15] = (yy - xx) * T;
        } else if (sh == 1) {
            w[0] = 2.0 / 3.0; w[1] = x * (1.0 / (3.0 * c)); w[2] = y * (1.0 / (3.0 * c)); w[3] = B;
            w[4] = x * y * F; w[5] = I; w[6] = I; w[7] = -x * (Nn / 2); w[8] = -y * (Nn / 2);
            w[9] = x * P; w[10] = y * P; w[11] = y * (Nn / 2); w[12] = x * (Nn / 2);
            w[13] = -x * y * S; w[14] = S / 2; w[15] = S / 2;
        } else if (sh == 2) {
            w[0] = -0.5; w[3] = D; w[5] = xx * J + yy * K; w[6] = xx * K + yy * J;
            w[7] = -x * (Nn / 8); w[8] = -y * (Nn / 8); w[9] = x * Qq; w[10] = y * Qq;
            w[11] = y * R; w[12] = x * R; w[14] = xx * U + yy * V; w[15] = xx * V + yy * U;
        } else {
            w[0] = 1.0 / 12; w[1] = -x * (1.0 / (24 * c)); w[2] = -y * (1.0 / (24 * c)); w[3] = E;
            w[4] = -x * y * (F / 16); w[5] = L; w[6] = L; w[7] = x * (Nn / 16); w[8] = y * (Nn / 16);
            w[9] = -x * (P / 8); w[10] = -y * (P / 8); w[11] = -y * (R / 2); w[12] = -x * (R / 2);
            w[13] = x * y * (S / 4); w[14] = S / 4; w[15] = S / 4;
        }
    }
}

/* test access to the D2V16 tables */
void ora_d2v16_tables(double *a, double *xi_x, double *xi_y, double *e)
{
    double t[QMAX][QMAX];
    d2v16_tables(t, xi_x, xi_y, e);
    if (a) memcpy(a, t, sizeof(t));
}

/* Pseudopotential equilibrium of population q.  D2Q9: Ark/Lib/D2Q9PP.cpp:51-62.
 * D2V16: f_eq_q = sum_k iArkC[q][k] M16_k (Update_phi_Eq, D2V16.cpp:157-166)
 * with Macro16 (:125-145) at the reference temperature, p = rho R0 T0. */
static double pp_feq(const ora *o, int q, double rho, double U, double V)
{
    const double rt = o->c.rt;
    if (o->Q == Q9) {
        double ex = o->xi[0][q], ey = o->xi[1][q], uu = U * U + V * V;
        double u1 = (ex * U + ey * V) / rt;
        return W9[q] * rho * (1.0 + u1 + 0.5 * u1 * u1 - uu * o->lambda0);
    }
    double M[16], u2 = U * U, v2 = V * V, p = rho * rt;
    const double deg = 3.0 + 0, degp2 = deg + 2.0;
    M[0] = rho;
    M[1] = rho * U;
    M[2] = rho * V;
    M[3] = rho * (deg * rt + u2 + v2);
    M[4] = M[1] * V;
    M[5] = rho * u2 + p;
    M[6] = rho * v2 + p;
    M[7] = M[1] * (degp2 * rt + u2 + v2);
    M[8] = M[2] * (degp2 * rt + u2 + v2);
    M[9] = M[1] * (3.0 * rt + u2);
    M[10] = M[2] * (3.0 * rt + v2);
    M[11] = M[2] * (rt + u2);
    M[12] = M[1] * (rt + v2);
    M[13] = ((deg + 4.0) * rt + u2 + v2) * M[4];
    M[14] = degp2 * rho * rt * rt + p * ((deg + 5.0) * u2 + v2) + rho * u2 * (u2 + v2);
    M[15] = degp2 * rho * rt * rt + p * ((deg + 5.0) * v2 + u2) + rho * v2 * (u2 + v2);
    double sum = 0;
    for (int k = 0; k < 16; ++k) sum += o->v16a[q][k] * M[k];
    return sum;
}

/* test access: D2V16 equilibrium at (rho, U, V) and temperature rt */
void ora_d2v16_feq(double rho, double U, double V, double rt, double *feq)
{
    ora o;
    memset(&o, 0, sizeof(o));
    o.Q = 16;
    o.c.rt = rt;
    d2v16_tables(o.v16a, o.xi[0], o.xi[1], NULL);
    for (int q = 0; q < 16; ++q) feq[q] = pp_feq(&o, q, rho, U, V);
}

/* ---------------- construction ---------------- */

ora *ora_create(const mpd_config *cfg)
{
    if (!cfg || cfg->nx < 3 || cfg->ny < 3) return NULL;
    if (cfg->model != MPD_MODEL_AC_P && cfg->model != MPD_MODEL_AC_HCZ &&
        cfg->model != MPD_MODEL_PP_D2Q9 && cfg->model != MPD_MODEL_PP_D2V16)
        return NULL;
    if (cfg->model == MPD_MODEL_PP_D2V16 && (cfg->flux != MPD_FLUX_CD || cfg->bc_y != MPD_BC_PERIODIC))
        return NULL; /* D2V16: CD faces on a periodic mesh */
    if (cfg->bc_y != MPD_BC_PERIODIC && cfg->model == MPD_MODEL_PP_D2Q9)
        return NULL; /* the reference has no working wall path for it */
    ora *o = (ora *)calloc(1, sizeof(ora));
    o->c = *cfg;
    o->nx = cfg->nx;
    o->ny = cfg->ny;
    o->N = cfg->nx * cfg->ny;
    o->walls = cfg->bc_y == MPD_BC_WALL_BB || cfg->bc_y == MPD_BC_WALL_NEE;
    o->nyf = o->walls ? o->ny + 1 : o->ny;
    const mpd_config *c = cfg;
    o->hdt = 0.5 * c->dt;
    o->lambda0 = 1 / (2.0 * c->rt);
    o->aphi = (c->rho_l - c->rho_v) / (c->phi_l - c->phi_v);
    o->bphi = c->rho_v - o->aphi * c->phi_v;
    o->beta = 12.0 * c->sigma / c->width;
    o->kappa = 3.0 * c->sigma * c->width / 2;
    o->kp = c->rt * 9 / 5.0;
    o->taum = c->mobility / c->rt;
    o->pp = c->model == MPD_MODEL_PP_D2Q9 || c->model == MPD_MODEL_PP_D2V16;
    if (c->model == MPD_MODEL_PP_D2V16) {
        o->Q = 16;
        d2v16_tables(o->v16a, o->xi[0], o->xi[1], NULL);
    } else {
        o->Q = Q9;
        double maspan = sqrt(3.0 * c->rt);
        for (int q = 0; q < Q9; ++q) {
            o->xi[0][q] = maspan * E9X[q];
            o->xi[1][q] = maspan * E9Y[q];
        }
    }
    o->dtv = c->dt / (c->dx * c->dx);
    /* Face_2D::DVDF::setxh / Cell_2D::DVDF::setxBP with tau_h = TauMass,
     * set once at initialisation (Ark/Cpp/Preprocess.cpp:727-734,751-758). */
    o->ah_h = 2.0 * o->taum / (2.0 * o->taum + o->hdt);
    o->bh_h = 1.0 - o->ah_h;
    o->ch_h = o->taum * o->bh_h;
    o->abp_h = (2.0 * o->taum - o->hdt) / (2.0 * o->taum + c->dt);
    o->bbp_h = 1.0 - o->abp_h;
    o->cbp_h = o->taum * o->bbp_h;
    o->ah_pp = 2.0 * c->pp_tau / (2.0 * c->pp_tau + o->hdt);
    o->bh_pp = 1.0 - o->ah_pp;
    o->ch_pp = c->pp_tau * o->bh_pp;
    size_t N = o->N, QN = (size_t)o->Q * N;
    o->h = calloc(QN, sizeof(double));
    o->f = calloc(QN, sizeof(double));
    o->m = calloc((size_t)MPD_M_COUNT * N, sizeof(double));
    o->ffx = calloc(2 * N, sizeof(double));
    o->ffy = calloc(2 * N, sizeof(double));
    o->fpsi = calloc(2 * N, sizeof(double));
    o->hbp = calloc(QN, sizeof(double));
    o->fbp = calloc(QN, sizeof(double));
    o->hbpx = calloc(QN, sizeof(double));
    o->hbpy = calloc(QN, sizeof(double));
    o->fbpx = calloc(QN, sizeof(double));
    o->fbpy = calloc(QN, sizeof(double));
    o->hfl = calloc(QN + (size_t)o->Q * o->nyf * o->nx, sizeof(double));
    o->ffl = calloc(QN + (size_t)o->Q * o->nyf * o->nx, sizeof(double));
    o->gm = calloc((size_t)MPD_M_COUNT * 2 * o->nx, sizeof(double));
    o->ghbp = calloc((size_t)Q9 * 2 * o->nx, sizeof(double));
    o->gfbp = calloc((size_t)Q9 * 2 * o->nx, sizeof(double));
    o->fpsi_old = calloc(2 * N, sizeof(double));
    o->cbfx = calloc((size_t)o->nx + o->ny, sizeof(double));
    o->cbfy = calloc((size_t)o->nx + o->ny, sizeof(double));
    o->cbfl = calloc((size_t)o->Q * (o->nx + o->ny), sizeof(double));
    o->cbhl = calloc((size_t)o->Q * (o->nx + o->ny), sizeof(double));
    o->cbpsi = calloc((size_t)o->nx + o->ny, sizeof(double));
    o->cbpsi_old = calloc((size_t)o->nx + o->ny, sizeof(double));
    return o;
}

void ora_destroy(ora *o)
{
    if (!o) return;
    double *bufs[] = {o->h, o->f, o->m, o->ffx, o->ffy, o->fpsi, o->hbp, o->fbp, o->hbpx,
                      o->hbpy, o->fbpx, o->fbpy, o->hfl, o->ffl, o->gm, o->ghbp, o->gfbp,
                      o->fpsi_old, o->cbfx, o->cbfy, o->cbfl, o->cbhl, o->cbpsi, o->cbpsi_old};
    for (size_t k = 0; k < sizeof(bufs) / sizeof(bufs[0]); ++k) free(bufs[k]);
    free(o);
}

int ora_set_state(ora *o, const double *h, const double *f, const double *m,
                  const double *face_fx, const double *face_fy)
{
    size_t N = o->N, QN = (size_t)o->Q * N;
    if (h) memcpy(o->h, h, QN * sizeof(double));
    if (f) memcpy(o->f, f, QN * sizeof(double));
    if (m) memcpy(o->m, m, (size_t)MPD_M_COUNT * N * sizeof(double));
    if (face_fx) memcpy(o->ffx, face_fx, 2 * N * sizeof(double));
    if (face_fy) memcpy(o->ffy, face_fy, 2 * N * sizeof(double));
    /* both stored copies of a periodic face start from the uploaded force */
    for (int j = 0; j < o->ny; ++j) {
        o->cbfx[j] = o->ffx[(size_t)j * o->nx];
        o->cbfy[j] = o->ffy[(size_t)j * o->nx];
    }
    for (int i = 0; i < o->nx; ++i) {
        o->cbfx[o->ny + i] = o->ffx[N + i];
        o->cbfy[o->ny + i] = o->ffy[N + i];
    }
    return 0;
}

int ora_set_face_order(ora *o, int ref_order)
{
    if (ref_order && !(o->c.model == MPD_MODEL_PP_D2Q9 && !o->walls))
        return -1;
    o->ref_order = ref_order != 0;
    return 0;
}

/* Test knob: Allen-Cahn UW with one face per location (the last column and
 * row take the first copy's flux), i.e. without the reference's second copies. */
int ora_set_one_copy(ora *o, int one)
{
    o->one_copy = one != 0;
    return 0;
}

int ora_set_face_extra(ora *o, const double *psi_old, const double *copy_fx, const double *copy_fy)
{
    if (psi_old) {
        memcpy(o->fpsi_old, psi_old, 2 * (size_t)o->N * sizeof(double));
        /* the second copies take the first copies' previous Psi (the same bits with CD faces) */
        for (int j = 0; j < o->ny; ++j) o->cbpsi_old[j] = psi_old[(size_t)j * o->nx];
        for (int i = 0; i < o->nx; ++i) o->cbpsi_old[o->ny + i] = psi_old[o->N + i];
    }
    if (copy_fx) memcpy(o->cbfx, copy_fx, ((size_t)o->nx + o->ny) * sizeof(double));
    if (copy_fy) memcpy(o->cbfy, copy_fy, ((size_t)o->nx + o->ny) * sizeof(double));
    return 0;
}

int ora_get_face_extra(const ora *o, double *psi, double *copy_fx, double *copy_fy)
{
    if (psi) memcpy(psi, o->fpsi_old, 2 * (size_t)o->N * sizeof(double));
    if (copy_fx) memcpy(copy_fx, o->cbfx, ((size_t)o->nx + o->ny) * sizeof(double));
    if (copy_fy) memcpy(copy_fy, o->cbfy, ((size_t)o->nx + o->ny) * sizeof(double));
    return 0;
}

int ora_get_state(const ora *o, double *h, double *f, double *m, double *face_fx,
                  double *face_fy)
{
    size_t N = o->N, QN = (size_t)o->Q * N;
    if (h) memcpy(h, o->h, QN * sizeof(double));
    if (f) memcpy(f, o->f, QN * sizeof(double));
    if (m) memcpy(m, o->m, (size_t)MPD_M_COUNT * N * sizeof(double));
    if (face_fx) memcpy(face_fx, o->ffx, 2 * N * sizeof(double));
    if (face_fy) memcpy(face_fy, o->ffy, 2 * N * sizeof(double));
    return 0;
}

/* ---------------- phase 1: cell equilibria, sources, bar-plus ---------------- */

/* Update_phi_Eq + Update_DVDF_Source + Update_phi_BP for one cell.
 * HCZ: Ark/Lib/D2Q9HCZ.cpp:54-92,244-257; AC-p: Ark/Lib/D2Q9AC.cpp:69-105,331-344;
 * PP: Ark/Lib/D2Q9PP.cpp:51-62,137-145; bar-plus: Ark/Cpp/DUGKS2DSolver.cpp:379-401;
 * factors: Ark/Cpp/Mesh/Cell_2D.cpp:120-125. */
static void phase1_cell(ora *o, size_t n)
{
    const double rt = o->c.rt;
    double *m = o->m;
    size_t N = o->N;
    double phi = m[MPD_M_PHI * N + n], rho = m[MPD_M_RHO * N + n];
    double U = m[MPD_M_U * N + n], V = m[MPD_M_V * N + n], p = m[MPD_M_P * N + n];
    double Fx = m[MPD_M_FX * N + n], Fy = m[MPD_M_FY * N + n];
    double rx = m[MPD_M_RHOX * N + n], ry = m[MPD_M_RHOY * N + n];
    double px = m[MPD_M_PHIX * N + n], py = m[MPD_M_PHIY * N + n];
    double tau = m[MPD_M_TAU * N + n];
    double abp = (2.0 * tau - o->hdt) / (2.0 * tau + o->c.dt);
    double bbp = 1.0 - abp;
    double cbp = tau * bbp;
    int model = o->c.model;
    double uu = U * U + V * V;
    for (int q = 0; q < o->Q; ++q) {
        double ex = o->xi[0][q], ey = o->xi[1][q], w = o->pp ? 0.0 : W9[q];
        size_t k = (size_t)q * N + n;
        double feq, fso;
        if (o->pp) {
            feq = pp_feq(o, q, rho, U, V);
            fso = (Fx * (ex - U) + Fy * (ey - V)) * feq / (rho * rt);
        } else {
            double u1 = (ex * U + ey * V) / rt;
            double G = u1 + 0.5 * u1 * u1 - uu * o->lambda0;
            double heq = w * (1.0 + u1) * phi;
            double hso = (px * ex + py * ey) * w;
            if (model == MPD_MODEL_AC_HCZ) {
                feq = w * (p + rho * G * rt);
                fso = w * ((1 + G) * ((ex - U) * Fx + (ey - V) * Fy) +
                           G * rt * ((ex - U) * rx + (ey - V) * ry));
            } else {
                feq = w * (p / rt + rho * G) - (double)(q == 0) * p / rt;
                fso = w * ((1 + G) * ((ex - U) * Fx + (ey - V) * Fy) / rt +
                           G * ((ex - U) * rx + (ey - V) * ry));
            }
            o->hbp[k] = o->abp_h * o->h[k] + o->bbp_h * heq + o->cbp_h * hso;
        }
        double b = abp * o->f[k] + bbp * feq;
        b += cbp * fso;
        o->fbp[k] = b;
    }
}

/* ---------------- phase 2: wall ghosts ---------------- */

/* WallShadowC_fBP / ExtrapolationfBP, Ark/Cpp/Boundary.cpp:38-85: ghost macros are
 * 2*near - next (MacroQuantity arithmetic, Ark/Cpp/Mesh/Macroscopic.cpp:20-76, which
 * leaves scG = next.scG), with U, V, Mu replaced by -near; bar-plus extrapolated. */
static void phase2_walls(ora *o)
{
    int nx = o->nx, ny = o->ny;
    for (int side = 0; side < 2; ++side) {
        int jn = side == 0 ? 0 : ny - 1;
        int jnn = side == 0 ? 1 : ny - 2;
        for (int i = 0; i < nx; ++i) {
            for (int fld = 0; fld < MPD_M_COUNT; ++fld) {
                double near = M(o, fld, i, jn), next = M(o, fld, i, jnn), g;
                if (fld == MPD_M_U || fld == MPD_M_V || fld == MPD_M_MU) g = -near;
                else if (fld == MPD_M_SCG) g = next;
                else g = 2.0 * near - next;
                o->gm[((size_t)fld * 2 + side) * nx + i] = g;
            }
            for (int q = 0; q < Q9; ++q) {
                size_t a = (size_t)q * o->N + (size_t)jn * nx + i;
                size_t b = (size_t)q * o->N + (size_t)jnn * nx + i;
                o->ghbp[((size_t)q * 2 + side) * nx + i] = 2 * o->hbp[a] - o->hbp[b];
                o->gfbp[((size_t)q * 2 + side) * nx + i] = 2 * o->fbp[a] - o->fbp[b];
            }
        }
    }
}

/* ---------------- phase 3: UW gradients ---------------- */

/* update_DVDF_Grad6points, Ark/Lib/GradSchemeBasic.h:95-126 (apsi = 4, bpsi = 1).
 * With walls, a wall-row cell's stencil reaches the wall's shadow cells, which
 * hold the extrapolated bar-plus of phase 2; its diagonal neighbour across the
 * periodic x line is PeriodicShadowC_SE/SW/NE/NW (DiagonalCellConstruct,
 * Ark/Cpp/Mesh/MeshCartesian.cpp:61-118), an alias of the shadow cell of the
 * wall row's other end (ShadowCellAssign shares the population arrays,
 * Ark/Cpp/Mesh/GhostCells.cpp:20-28,38-64), i.e. the ghost row wrapped in x. */
static void phase3_grad(ora *o, const double *bp, const double *gbp, double *gx, double *gy)
{
    int nx = o->nx, ny = o->ny;
    double d2x = 2.0 * o->c.dx;
    for (int q = 0; q < Q9; ++q)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                double C0 = BP(o, bp, gbp, q, i + 1, j), C1 = BP(o, bp, gbp, q, i, j + 1);
                double C2 = BP(o, bp, gbp, q, i - 1, j), C3 = BP(o, bp, gbp, q, i, j - 1);
                double D0 = BP(o, bp, gbp, q, i + 1, j + 1), D1 = BP(o, bp, gbp, q, i - 1, j + 1);
                double D2 = BP(o, bp, gbp, q, i - 1, j - 1), D3 = BP(o, bp, gbp, q, i + 1, j - 1);
                size_t k = (size_t)q * o->N + (size_t)j * nx + i;
                gx[k] = (4 * (C0 - C2) + 1 * (D0 - D1 + D3 - D2)) / (d2x * 6);
                gy[k] = (4 * (C1 - C3) + 1 * (D0 - D3 + D1 - D2)) / (d2x * 6);
            }
}

/* ---------------- phase 4: faces ---------------- */

typedef struct {
    int dir;        /* 0: x-face (normal +x), 1: y-face (normal +y) */
    int oi, oj;     /* owner cell */
    int ni, nj;     /* neighbour cell */
    double xf, yf;  /* face centre */
} face_t;

/* f-bar at the face, t + h: CD_Interior_phi_Bh (Ark/Cpp/DUGKS2DSolver.cpp:604-687)
 * or UW_Interior_phi_Bh (:560-603, selection :688-706).
 *
 * CD, x-face: the reference's normal derivative (owner - neigh)/_dx with
 * _dx = xc_owner - xc_neigh (MeshCartesian.cpp:419-427) equals (n - o)/dx here,
 * and its tangential derivative (0.5*(fc3-fc1) + 0.5*(fc2-fc0))/_dy with
 * _dy = +-2dx (MeshCartesian.cpp:145,203) equals the two half-differences of
 * the owner and neighbour columns over 2dx; all sign changes are exact. */
static double face_recon(const ora *o, const double *bp, const double *gbp, const double *gx,
                         const double *gy, const face_t *F, int q)
{
    double ex = o->xi[0][q], ey = o->xi[1][q], dx = o->c.dx;
    double bo = BP(o, bp, gbp, q, F->oi, F->oj), bn = BP(o, bp, gbp, q, F->ni, F->nj);
    if (o->c.flux == MPD_FLUX_CD) {
        double dn = (bn - bo) / dx, dt;
        if (F->dir == 0) {
            double No = BP(o, bp, gbp, q, F->oi, F->oj + 1), So = BP(o, bp, gbp, q, F->oi, F->oj - 1);
            double Nn = BP(o, bp, gbp, q, F->ni, F->nj + 1), Sn = BP(o, bp, gbp, q, F->ni, F->nj - 1);
            dt = (0.5 * (Nn - Sn) + 0.5 * (No - So)) / (2 * dx);
            return 0.5 * (bo + bn) - o->hdt * (dn * ex + dt * ey);
        } else {
            double Eo = BP(o, bp, gbp, q, F->oi + 1, F->oj), Wo = BP(o, bp, gbp, q, F->oi - 1, F->oj);
            double En = BP(o, bp, gbp, q, F->ni + 1, F->nj), Wn = BP(o, bp, gbp, q, F->ni - 1, F->nj);
            dt = (0.5 * (En - Wn) + 0.5 * (Eo - Wo)) / (2 * dx);
            return 0.5 * (bo + bn) - o->hdt * (dt * ex + dn * ey);
        }
    }
    /* UW: upwind along xi.n of the face's own normal */
    double xin = F->dir == 0 ? ex : ey;
    double xo = (F->oi + 0.5) * dx, yo = (F->oj + 0.5) * dx;
    double xn = (F->ni + 0.5) * dx, yn = (F->nj + 0.5) * dx;
    size_t ko = (size_t)q * o->N + (size_t)wrap(F->oj, o->ny) * o->nx + wrap(F->oi, o->nx);
    size_t kn = (size_t)q * o->N + (size_t)wrap(F->nj, o->ny) * o->nx + wrap(F->ni, o->nx);
    double dxo = F->xf - o->hdt * ex - xo, dyo = F->yf - o->hdt * ey - yo;
    double dxn = F->xf - o->hdt * ex - xn, dyn = F->yf - o->hdt * ey - yn;
    /* A wall face's other side is the wall's shadow cell.  Its gradient is
     * never computed (Grad_VS_6points runs over the mesh cells, :197-206) and
     * its arrays come from `new double[]` (AllocateARK, Ark/Cpp/Mesh/Mesh_2D.cpp:8-13),
     * so the reference reads whatever that memory holds: 0 in the compiled
     * reference's runs, and 0 here.  It enters the face macros through the
     * populations that Wall_3_BB then replaces and through the xi.n = 0 average. */
    double gxo = 0.0, gyo = 0.0, gxn = 0.0, gyn = 0.0;
    if (!(o->walls && F->oj < 0)) { gxo = gx[ko]; gyo = gy[ko]; }
    if (!(o->walls && F->nj >= o->ny)) { gxn = gx[kn]; gyn = gy[kn]; }
    double ro = bo + (dxo * gxo + dyo * gyo), rn = bn + (dxn * gxn + dyn * gyn);
    if (xin > 0) return ro;
    if (xin < 0) return rn;
    /* xi.n = 0: the average adds the reference face's owner first.  A boundary
     * face's owner is its mesh cell (ConstructFacesCells,
     * Ark/Lib/MeshConstructFunc.h:216-233): at index 0 of its normal (the
     * zone-12 copy of a periodic face, the bottom wall) that is this face's
     * neighbour; the zone-8 copy (make_face at i = nx / j = ny) has it as owner. */
    if (F->oi < 0 || F->oj < 0) return (bn + (dxn * gxn + dyn * gyn) + bo + (dxo * gxo + dyo * gyo)) / 2;
    return (bo + (dxo * gxo + dyo * gyo) + bn + (dxn * gxn + dyn * gyn)) / 2;
}

typedef struct {
    double phi, rho, U, V, p, tau, px, py, rx, ry, Fx, Fy, scg, psi;
} face_macro;

/* Update_MacroVar_h, HCZ Ark/Lib/D2Q9HCZ.cpp:183-242, AC-p Ark/Lib/D2Q9AC.cpp:232-329,
 * PP Ark/Lib/D2Q9PP.cpp:110-136 (+ Update_PseudoPsi(face), Ark/Cpp/PseudoP.cpp:23-30);
 * moment sums in sequential order as IntegralGH (Ark/Lib/GaussHermite.h:1-14). */
static void face_macros(const ora *o, const face_t *F, const double *hb, const double *fb,
                        double fx_old, double fy_old, face_macro *fm)
{
    const double rt = o->c.rt, hdt = o->hdt;
    int model = o->c.model;
    (void)model;
    if (o->pp) {
        double s = 0, sx = 0, sy = 0;
        for (int q = 0; q < o->Q; ++q) s += fb[q];
        for (int q = 0; q < o->Q; ++q) sx += fb[q] * o->xi[0][q];
        for (int q = 0; q < o->Q; ++q) sy += fb[q] * o->xi[1][q];
        fm->rho = s;
        fm->U = sx / fm->rho;
        fm->V = sy / fm->rho;
        fm->U += 0.5 * hdt * fx_old / fm->rho;
        fm->V += 0.5 * hdt * fy_old / fm->rho;
        fm->p = o->c.pp_k * carnahan_starling(o, fm->rho);
        calc_psi(fm->p, fm->rho, rt, &fm->scg, &fm->psi);
        fm->tau = o->c.pp_tau;
        return;
    }
    fm->px = 0.5 * (M(o, MPD_M_PHIX, F->oi, F->oj) + M(o, MPD_M_PHIX, F->ni, F->nj));
    fm->py = 0.5 * (M(o, MPD_M_PHIY, F->oi, F->oj) + M(o, MPD_M_PHIY, F->ni, F->nj));
    fm->rx = 0.5 * (M(o, MPD_M_RHOX, F->oi, F->oj) + M(o, MPD_M_RHOX, F->ni, F->nj));
    fm->ry = 0.5 * (M(o, MPD_M_RHOY, F->oi, F->oj) + M(o, MPD_M_RHOY, F->ni, F->nj));
    fm->Fx = 0.5 * (M(o, MPD_M_FX, F->oi, F->oj) + M(o, MPD_M_FX, F->ni, F->nj));
    fm->Fy = 0.5 * (M(o, MPD_M_FY, F->oi, F->oj) + M(o, MPD_M_FY, F->ni, F->nj));
    double sh = 0;
    for (int q = 0; q < Q9; ++q) sh += hb[q];
    fm->phi = sh;
    fm->rho = o->aphi * fm->phi + o->bphi;
    double s = 0, sx = 0, sy = 0;
    for (int q = 0; q < Q9; ++q) s += fb[q];
    double mu;
    if (model == MPD_MODEL_AC_HCZ) {
        for (int q = 0; q < Q9; ++q) sx += fb[q] * o->xi[0][q];
        for (int q = 0; q < Q9; ++q) sy += fb[q] * o->xi[1][q];
        fm->U = sx / (fm->rho * rt);
        fm->V = sy / (fm->rho * rt);
        fm->U += 0.5 * hdt * fm->Fx / (fm->rho * rt);
        fm->V += 0.5 * hdt * fm->Fy / (fm->rho * rt);
        fm->p = s + 0.5 * hdt * rt * (fm->rx * fm->U + fm->ry * fm->V);
        mu = o->c.nu_v + (fm->phi - o->c.phi_v) * (o->c.nu_l - o->c.nu_v);
        mu *= fm->rho;
    } else {
        fm->U = 0.0;
        fm->V = 0.0;
        fm->p = s - fb[0];
        fm->p += (0.5 * hdt * (fm->rx * fm->U + fm->ry * fm->V) -
                  W9[0] * fm->rho * (fm->U * fm->U + fm->V * fm->V) * o->lambda0);
        fm->p *= o->kp;
        mu = fm->rho * rt * o->c.tau0;
    }
    fm->tau = mu / (fm->rho * rt);
}

/* Update_phi_Eqh + Update_DVDF_Source_h + Update_phi_h for one population.
 * HCZ Ark/Lib/D2Q9HCZ.cpp:93-131,258-271; AC-p Ark/Lib/D2Q9AC.cpp:106-144,345-358;
 * PP Ark/Lib/D2Q9PP.cpp:63-74,146-154; BGK half step Ark/Cpp/DUGKS2DSolver.cpp:402-423;
 * factors Ark/Cpp/Mesh/Face_2D.cpp:123-128. */
static void face_bgk(const ora *o, const face_macro *fm, int q, double hb, double fb,
                     double ah, double bh, double ch, double *hout, double *fout)
{
    const double rt = o->c.rt;
    double ex = o->xi[0][q], ey = o->xi[1][q], w = o->pp ? 0.0 : W9[q];
    double U = fm->U, V = fm->V, uu = U * U + V * V;
    double u1 = (ex * U + ey * V) / rt;
    double feq, fso;
    if (o->pp) {
        feq = pp_feq(o, q, fm->rho, U, V);
        fso = (fm->Fx * (ex - U) + fm->Fy * (ey - V)) * feq / (fm->rho * rt);
        *hout = 0;
    } else {
        double G = u1 + 0.5 * u1 * u1 - uu * o->lambda0;
        double heq = w * fm->phi * (1.0 + u1);
        double hso = (fm->px * ex + fm->py * ey) * w;
        *hout = o->ah_h * hb + o->bh_h * heq + o->ch_h * hso;
        if (o->c.model == MPD_MODEL_AC_HCZ) {
            feq = w * (fm->p + fm->rho * rt * G);
            fso = w * ((1 + G) * ((ex - U) * fm->Fx + (ey - V) * fm->Fy) +
                       G * rt * ((ex - U) * fm->rx + (ey - V) * fm->ry));
        } else {
            feq = w * (fm->p / rt + fm->rho * G) - (double)(q == 0) * fm->p / rt;
            fso = w * ((1 + G) * ((ex - U) * fm->Fx + (ey - V) * fm->Fy) / rt +
                       G * ((ex - U) * fm->rx + (ey - V) * fm->ry));
        }
    }
    double r = ah * fb + bh * feq;
    r += ch * fso;
    *fout = r;
}

static void make_face(const ora *o, int dir, int i, int j, face_t *F)
{
    F->dir = dir;
    if (dir == 0) {
        F->oi = i - 1; F->oj = j; F->ni = i; F->nj = j;
        F->xf = i * o->c.dx; F->yf = (j + 0.5) * o->c.dx;
    } else {
        F->oi = i; F->oj = j - 1; F->ni = i; F->nj = j;
        F->xf = (i + 0.5) * o->c.dx; F->yf = j * o->c.dx;
    }
}

/* psi of the face lattice of direction `dir` at (i,j), periodic in both directions */
static inline double face_psi(const ora *o, int dir, int i, int j)
{
    return o->fpsi[(size_t)dir * o->N + (size_t)wrap(j, o->ny) * o->nx + wrap(i, o->nx)];
}

/* ---- pseudopotential faces in the reference's one-thread order ----
 *
 * Face objects.  An x-face object is (i, j) with i in [0, nx], j in [0, ny):
 * i = 0 and i = nx are the two stored copies of the periodic face between
 * columns nx-1 and 0 (zone 12, owner (0,j); zone 8, owner (nx-1,j)); y-face
 * objects likewise with j in [0, ny].
 *
 * Rank: the position of the object in the visit of one thread, i.e. its index
 * in FaceArray (the interior loop runs over InteriorFaceA, then the periodic
 * loop over PeriodicFaceA, Ark/Cpp/DUGKS2DSolver.cpp:219-230, both in FaceArray
 * order).  FaceArray follows the face sections of the mesh file; for the mesh
 * written by mpdugks_amd.mesh.write_cas that is: interior x-faces row by row,
 * interior y-faces row by row, then the left (12), right (8), bottom (12) and
 * top (8) periodic zones. */
static long face_rank(const ora *o, int dir, int i, int j)
{
    long nx = o->nx, ny = o->ny, NX = ny * (nx - 1), NI = NX + (ny - 1) * nx;
    if (dir == 0) {
        if (i == 0) return NI + j;
        if (i == nx) return NI + ny + j;
        return (long)j * (nx - 1) + (i - 1);
    }
    if (j == 0) return NI + 2 * ny + i;
    if (j == ny) return NI + 2 * ny + nx + i;
    return NX + (long)(j - 1) * nx + i;
}

/* The object a face stencil reaches from object (dir, *i, *j) at offset
 * (di, dj).  faceFaces (Ark/Cpp/Mesh/MeshCartesian.cpp:296-368) take a periodic
 * face through the cells' own Face_C, or through DiagonalFacesCheck (:286-294)
 * for the diagonal entries; either way the copy on the referencing face's side
 * of the domain, the geometrically nearer one: one step across the periodic
 * line from i = 0 lands on i = nx - 1, from i = nx on 1. */
static void face_obj_step(const ora *o, int dir, int *i, int *j, int di, int dj)
{
    int nx = o->nx, ny = o->ny, a = *i + di, b = *j + dj;
    if (dir == 0) {
        if (a < 0) a += nx;
        else if (a > nx) a -= nx;
        b = wrap(b, ny);
    } else {
        a = wrap(a, nx);
        if (b < 0) b += ny;
        else if (b > ny) b -= ny;
    }
    *i = a;
    *j = b;
}

/* Psi of the neighbour at (di, dj) as object `self` (rank r0) reads it: the new
 * value if one thread has already visited it this step, else the value it kept
 * from the previous step (msqh->Psi, Ark/Lib/GradSchemeBasic.h:181-208). */
static double psi_seen(const ora *o, int dir, int i, int j, int di, int dj, long r0)
{
    face_obj_step(o, dir, &i, &j, di, dj);
    const int seen = face_rank(o, dir, i, j) < r0;
    if ((dir == 0 && i == o->nx) || (dir == 1 && j == o->ny)) {  /* a second copy */
        size_t b = dir == 0 ? (size_t)j : (size_t)o->ny + i;
        return seen ? o->cbpsi[b] : o->cbpsi_old[b];
    }
    size_t fi = (size_t)dir * o->N + (size_t)wrap(j, o->ny) * o->nx + wrap(i, o->nx);
    return seen ? o->fpsi[fi] : o->fpsi_old[fi];
}

/* Update_PseudoForce(Face_2D&) (Ark/Cpp/PseudoP.cpp:39-46) of object (dir, i, j)
 * with the face-lattice gradient update_MQ_x/y(Face_2D*) (GradSchemeBasic.h:181-208). */
static void pp_force_ref(const ora *o, int dir, int i, int j, double scg, double psi, double *fx, double *fy)
{
    long r0 = face_rank(o, dir, i, j);
    double d2x = 2.0 * o->c.dx;
    double F1 = psi_seen(o, dir, i, j, 1, 0, r0), F5 = psi_seen(o, dir, i, j, -1, 0, r0);
    double F3 = psi_seen(o, dir, i, j, 0, 1, r0), F7 = psi_seen(o, dir, i, j, 0, -1, r0);
    double F2 = psi_seen(o, dir, i, j, 1, 1, r0), F4 = psi_seen(o, dir, i, j, -1, 1, r0);
    double F6 = psi_seen(o, dir, i, j, -1, -1, r0), F8 = psi_seen(o, dir, i, j, 1, -1, r0);
    double gxp = (4 * (F1 - F5) + 1 * (F2 - F4 + F8 - F6)) / (d2x * 6);
    double gyp = (4 * (F3 - F7) + 1 * (F2 - F8 + F4 - F6)) / (d2x * 6);
    gxp *= -scg * psi;
    gyp *= -scg * psi;
    *fx = gxp;
    *fy = gyp;
}

/* Forces of every face object in reference order.  fms holds the macros of each
 * structured face with the force of its first copy, fmb those of the second
 * copies of the periodic faces (index as cbfx).  With CD faces only U, V differ
 * between the copies (they carry their own previous force, D2Q9PP.cpp:119-122)
 * and rho, scG and Psi are the same bits; with UW faces the copies'
 * reconstructions, and so all their macros, can differ in the last bits. */
static void pp_forces_ref(ora *o, face_macro *fms, face_macro *fmb)
{
    int nx = o->nx, ny = o->ny;
    size_t N = o->N;
    for (int dir = 0; dir < 2; ++dir)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t fi = (size_t)dir * N + (size_t)j * nx + i;
                face_macro *fm = &fms[fi];
                pp_force_ref(o, dir, i, j, fm->scg, fm->psi, &fm->Fx, &fm->Fy);
                o->ffx[fi] = fm->Fx;
                o->ffy[fi] = fm->Fy;
                if ((dir == 0 && i == 0) || (dir == 1 && j == 0)) {
                    size_t b = dir == 0 ? (size_t)j : (size_t)ny + i;
                    face_macro *mb = &fmb[b];
                    pp_force_ref(o, dir, dir == 0 ? nx : i, dir == 0 ? j : ny, mb->scg, mb->psi, &mb->Fx, &mb->Fy);
                    o->cbfx[b] = mb->Fx;
                    o->cbfy[b] = mb->Fy;
                }
            }
}

/* Equilibria and sources of population q of cell n (the AC part of
 * phase1_cell, same expressions). */
static void cell_eq_so(const ora *o, size_t n, int q, double *heq, double *feq, double *hso, double *fso)
{
    const double rt = o->c.rt;
    const double *m = o->m;
    size_t N = o->N;
    double phi = m[MPD_M_PHI * N + n], rho = m[MPD_M_RHO * N + n];
    double U = m[MPD_M_U * N + n], V = m[MPD_M_V * N + n], p = m[MPD_M_P * N + n];
    double Fx = m[MPD_M_FX * N + n], Fy = m[MPD_M_FY * N + n];
    double rx = m[MPD_M_RHOX * N + n], ry = m[MPD_M_RHOY * N + n];
    double px = m[MPD_M_PHIX * N + n], py = m[MPD_M_PHIY * N + n];
    double uu = U * U + V * V;
    double ex = o->xi[0][q], ey = o->xi[1][q], w = W9[q];
    double u1 = (ex * U + ey * V) / rt;
    double G = u1 + 0.5 * u1 * u1 - uu * o->lambda0;
    *heq = w * (1.0 + u1) * phi;
    *hso = (px * ex + py * ey) * w;
    if (o->c.model == MPD_MODEL_AC_HCZ) {
        *feq = w * (p + rho * G * rt);
        *fso = w * ((1 + G) * ((ex - U) * Fx + (ey - V) * Fy) + G * rt * ((ex - U) * rx + (ey - V) * ry));
    } else {
        *feq = w * (p / rt + rho * G) - (double)(q == 0) * p / rt;
        *fso = w * ((1 + G) * ((ex - U) * Fx + (ey - V) * Fy) / rt + G * ((ex - U) * rx + (ey - V) * ry));
    }
}

/* Wall_3_NEE (Ark/Cpp/Boundary.cpp:140-184): the wall face takes the
 * equilibrium of face macros extrapolated from its fluid cell (Phi and Rho
 * along their cell gradients -- Phi_x/Phi_y hold the AC source vector there --,
 * p of the cell, U = V = 0: the face macros are never set otherwise) plus the
 * cell's non-equilibrium part 2 tau / (2 tau + dt) (x~ - x_eq + h x_so); the
 * caller then applies Wall_3_BB. */
static void wall_nee(const ora *o, const face_t *F, int j, double *hh, double *ff)
{
    const double rt = o->c.rt, dx = o->c.dx;
    int jc = j == 0 ? 0 : o->ny - 1, i = F->oi;
    size_t N = o->N, n = (size_t)jc * o->nx + i;
    const double *m = o->m;
    double xc = (i + 0.5) * dx, yc = (jc + 0.5) * dx;
    double phi = m[MPD_M_PHI * N + n] + m[MPD_M_PHIX * N + n] * (F->xf - xc) + m[MPD_M_PHIY * N + n] * (F->yf - yc);
    double rho = m[MPD_M_RHO * N + n] + m[MPD_M_RHOX * N + n] * (F->xf - xc) + m[MPD_M_RHOY * N + n] * (F->yf - yc);
    double p = m[MPD_M_P * N + n], U = 0.0, V = 0.0, uu = U * U + V * V;
    double tau = m[MPD_M_TAU * N + n];
    double hNEq = 2.0 * o->taum / (2.0 * o->taum + o->c.dt);
    double fNEq = 2.0 * tau / (2.0 * tau + o->c.dt);
    for (int q = 0; q < Q9; ++q) {
        double heq, feq, hso, fso;
        cell_eq_so(o, n, q, &heq, &feq, &hso, &fso);
        /* Update_phi_Eqh with the face macros (expressions of face_bgk) */
        double ex = o->xi[0][q], ey = o->xi[1][q], w = W9[q];
        double u1 = (ex * U + ey * V) / rt;
        double G = u1 + 0.5 * u1 * u1 - uu * o->lambda0;
        double heqf = w * phi * (1.0 + u1), feqf;
        if (o->c.model == MPD_MODEL_AC_HCZ) feqf = w * (p + rho * rt * G);
        else feqf = w * (p / rt + rho * G) - (double)(q == 0) * p / rt;
        hh[q] = heqf + hNEq * (o->h[q * N + n] - heq + o->hdt * hso);
        ff[q] = feqf + fNEq * (o->f[q * N + n] - feq + o->hdt * fso);
    }
}

static void phase4_faces(ora *o)
{
    int nx = o->nx, ny = o->ny, N = o->N;
    int pp = o->pp, Q = o->Q;
    double hb[QMAX], fb[QMAX];
    face_macro *fms = NULL, *fmb = NULL;
    int ref = pp && o->ref_order;
    if (pp) fms = (face_macro *)malloc(sizeof(face_macro) * 2 * (size_t)N);
    if (ref) fmb = (face_macro *)malloc(sizeof(face_macro) * ((size_t)nx + ny));
    /* PP, first pass: face macros and Psi of every face (Jacobi order). */
    if (pp) {
        for (int dir = 0; dir < 2; ++dir)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    face_t F;
                    make_face(o, dir, i, j, &F);
                    for (int q = 0; q < Q; ++q)
                        fb[q] = face_recon(o, o->fbp, o->gfbp, o->fbpx, o->fbpy, &F, q);
                    size_t fi = (size_t)dir * N + (size_t)j * nx + i;
                    face_macro *fm = &fms[fi];
                    face_macros(o, &F, hb, fb, o->ffx[fi], o->ffy[fi], fm);
                    o->fpsi[fi] = fm->psi;
                    if (ref && ((dir == 0 && i == 0) || (dir == 1 && j == 0))) {
                        /* the second copy: its own reconstruction (with UW faces
                         * the xi.n = 0 average adds its own owner first), macros and Psi */
                        size_t b = dir == 0 ? (size_t)j : (size_t)ny + i;
                        face_t FB;
                        double fbb[QMAX];
                        make_face(o, dir, dir == 0 ? nx : i, dir == 0 ? j : ny, &FB);
                        for (int q = 0; q < Q; ++q)
                            fbb[q] = face_recon(o, o->fbp, o->gfbp, o->fbpx, o->fbpy, &FB, q);
                        face_macros(o, &FB, hb, fbb, o->cbfx[b], o->cbfy[b], &fmb[b]);
                        o->cbpsi[b] = fmb[b].psi;
                    }
                }
    }
    if (ref) pp_forces_ref(o, fms, fmb);
    else if (pp) {
        /* second pass: face force from the face-lattice gradient of Psi,
         * update_MQ_x/y(Face_2D*) (Ark/Lib/GradSchemeBasic.h:181-208) and
         * Update_PseudoForce(face) (Ark/Cpp/PseudoP.cpp:39-46) */
        double d2x = 2.0 * o->c.dx;
        for (int dir = 0; dir < 2; ++dir)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    size_t fi = (size_t)dir * N + (size_t)j * nx + i;
                    double F1 = face_psi(o, dir, i + 1, j), F5 = face_psi(o, dir, i - 1, j);
                    double F3 = face_psi(o, dir, i, j + 1), F7 = face_psi(o, dir, i, j - 1);
                    double F2 = face_psi(o, dir, i + 1, j + 1), F4 = face_psi(o, dir, i - 1, j + 1);
                    double F6 = face_psi(o, dir, i - 1, j - 1), F8 = face_psi(o, dir, i + 1, j - 1);
                    double gxp = (4 * (F1 - F5) + 1 * (F2 - F4 + F8 - F6)) / (d2x * 6);
                    double gyp = (4 * (F3 - F7) + 1 * (F2 - F8 + F4 - F6)) / (d2x * 6);
                    face_macro *fm = &fms[fi];
                    gxp *= -fm->scg * fm->psi;
                    gyp *= -fm->scg * fm->psi;
                    fm->Fx = gxp;
                    fm->Fy = gyp;
                    o->ffx[fi] = gxp;
                    o->ffy[fi] = gyp;
                }
    }
    for (int dir = 0; dir < 2; ++dir) {
        int rows = dir == 0 ? ny : o->nyf;
        double *hfl = o->hfl + (dir ? (size_t)Q * N : 0);
        double *ffl = o->ffl + (dir ? (size_t)Q * N : 0);
        size_t plane = dir == 0 ? (size_t)N : (size_t)o->nyf * nx;
        for (int j = 0; j < rows; ++j)
            for (int i = 0; i < nx; ++i) {
                face_t F;
                make_face(o, dir, i, j, &F);
                for (int q = 0; q < Q; ++q) {
                    if (!pp) hb[q] = face_recon(o, o->hbp, o->ghbp, o->hbpx, o->hbpy, &F, q);
                    fb[q] = face_recon(o, o->fbp, o->gfbp, o->fbpx, o->fbpy, &F, q);
                }
                face_macro fml, *fm;
                double ah, bh, ch;
                if (pp) {
                    fm = &fms[(size_t)dir * N + (size_t)j * nx + i];
                    ah = o->ah_pp; bh = o->bh_pp; ch = o->ch_pp;
                } else {
                    fm = &fml;
                    face_macros(o, &F, hb, fb, 0, 0, fm);
                    ah = 2.0 * fm->tau / (2.0 * fm->tau + o->hdt);
                    bh = 1.0 - ah;
                    ch = fm->tau * bh;
                }
                double hh[QMAX], ff[QMAX];
                for (int q = 0; q < Q; ++q)
                    face_bgk(o, fm, q, pp ? 0 : hb[q], fb[q], ah, bh, ch, &hh[q], &ff[q]);
                /* Wall_3_BB (Ark/Cpp/Boundary.cpp:118-139): populations entering the
                 * fluid take the value of their bounce-back partner. */
                int wall = o->walls && dir == 1 && (j == 0 || j == ny);
                if (wall && o->c.bc_y == MPD_BC_WALL_NEE) wall_nee(o, &F, j, hh, ff);
                if (wall) {
                    double h0[Q9], f0[Q9];
                    memcpy(h0, hh, sizeof(h0));
                    memcpy(f0, ff, sizeof(f0));
                    for (int q = 0; q < Q9; ++q) {
                        double ey = o->xi[1][q];
                        int incoming = (j == 0) ? (ey > 0) : (ey < 0);
                        if (incoming) { hh[q] = h0[BB9[q]]; ff[q] = f0[BB9[q]]; }
                    }
                }
                /* Update_phiFlux_h (:424-441): times xi.n dS (setXiDotdS,
                 * Ark/Cpp/Mesh/MeshConstruct.cpp:289-297), unit normal +x / +y */
                for (int q = 0; q < Q; ++q) {
                    double xin = o->c.dx * (dir == 0 ? o->xi[0][q] : o->xi[1][q]);
                    size_t k = (size_t)q * plane + (size_t)j * nx + i;
                    hfl[k] = hh[q] * xin;
                    ffl[k] = ff[q] * xin;
                }
                /* reference order: the second copy of a periodic face has its own
                 * macros and force, so its own flux (taken by the cell that owns it) */
                if (ref && ((dir == 0 && i == 0) || (dir == 1 && j == 0))) {
                    size_t b = dir == 0 ? (size_t)j : (size_t)ny + i;
                    face_t FB;
                    make_face(o, dir, dir == 0 ? nx : i, dir == 0 ? j : ny, &FB);
                    for (int q = 0; q < Q; ++q) {
                        double fbb = face_recon(o, o->fbp, o->gfbp, o->fbpx, o->fbpy, &FB, q);
                        face_bgk(o, &fmb[b], q, 0, fbb, ah, bh, ch, &hh[q], &ff[q]);
                        double xin = o->c.dx * (dir == 0 ? o->xi[0][q] : o->xi[1][q]);
                        o->cbfl[(size_t)q * (nx + ny) + b] = ff[q] * xin;
                    }
                }
            }
    }
    /* Allen-Cahn UW: the second stored copy (zone 8, owned by the last column /
     * row) of each periodic face, whose xi.n = 0 average adds the other cell
     * first, so its bits can differ from the first copy's. */
    if (!pp && o->c.flux == MPD_FLUX_UW && !o->one_copy) {
        int ncopy = o->walls ? ny : ny + nx;
        for (int b = 0; b < ncopy; ++b) {
            face_t F;
            if (b < ny) make_face(o, 0, nx, b, &F);
            else make_face(o, 1, b - ny, ny, &F);
            for (int q = 0; q < Q; ++q) {
                hb[q] = face_recon(o, o->hbp, o->ghbp, o->hbpx, o->hbpy, &F, q);
                fb[q] = face_recon(o, o->fbp, o->gfbp, o->fbpx, o->fbpy, &F, q);
            }
            face_macro fm;
            face_macros(o, &F, hb, fb, 0, 0, &fm);
            double ah = 2.0 * fm.tau / (2.0 * fm.tau + o->hdt), bh = 1.0 - ah, ch = fm.tau * bh;
            for (int q = 0; q < Q; ++q) {
                double hh, ff, xin = o->c.dx * (F.dir == 0 ? o->xi[0][q] : o->xi[1][q]);
                face_bgk(o, &fm, q, hb[q], fb[q], ah, bh, ch, &hh, &ff);
                o->cbhl[(size_t)q * (nx + ny) + b] = hh * xin;
                o->cbfl[(size_t)q * (nx + ny) + b] = ff * xin;
            }
        }
    }
    if (pp) memcpy(o->fpsi_old, o->fpsi, 2 * (size_t)N * sizeof(double));
    if (ref) memcpy(o->cbpsi_old, o->cbpsi, ((size_t)nx + ny) * sizeof(double));
    free(fms);
    free(fmb);
}

/* ---------------- phase 8: cell update ---------------- */

/* Update_phi_T, Ark/Cpp/DUGKS2DSolver.cpp:748-792, faces summed in the
 * reference's Face_C order E, N, W, S (sortFacesInThisCell,
 * Ark/Cpp/Mesh/MeshCartesian.cpp:28-60) with signFlux -1 for the owned E and N
 * faces and +1 for W and S (Ark/Cpp/Mesh/MeshConstruct.cpp:283-285). */
static void phase8_update(ora *o)
{
    int nx = o->nx, ny = o->ny, N = o->N;
    const double aTP = 4.0 / 3.0, bTP = 1.0 / 3.0;
    size_t yplane = (size_t)o->nyf * nx;
    int pp = o->pp, Q = o->Q;
    /* the last column / row owns the second copy of its periodic east / north
     * face (Face_C, MeshCartesian.cpp:28-60): its own flux in the pseudopotential
     * reference order and with Allen-Cahn UW faces */
    int copies = o->ref_order || (!pp && o->c.flux == MPD_FLUX_UW && !o->one_copy);
    size_t nb = (size_t)nx + ny;
    for (int q = 0; q < Q; ++q)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int cE = copies && i == nx - 1, cN = copies && !o->walls && j == ny - 1;
                size_t k = (size_t)q * N + (size_t)j * nx + i;
                size_t kE = (size_t)q * N + (size_t)j * nx + wrap(i + 1, nx);
                size_t kW = k;
                int jn = o->walls ? j + 1 : wrap(j + 1, ny);
                size_t kN = (size_t)q * yplane + (size_t)jn * nx + i;
                size_t kS = (size_t)q * yplane + (size_t)j * nx + i;
                const double *hx = o->hfl, *hy = o->hfl + (size_t)Q * N;
                const double *fx = o->ffl, *fy = o->ffl + (size_t)Q * N;
                if (!pp) {
                    double s = 0.0;
                    s += -1 * (cE ? o->cbhl[q * nb + j] : hx[kE]);
                    s += -1 * (cN ? o->cbhl[q * nb + ny + i] : hy[kN]);
                    s += 1 * hx[kW];
                    s += 1 * hy[kS];
                    o->h[k] = aTP * o->hbp[k] - bTP * o->h[k] + o->dtv * s;
                }
                double fE = cE ? o->cbfl[q * nb + j] : fx[kE];
                double fN = cN ? o->cbfl[q * nb + ny + i] : fy[kN];
                double s = 0.0;
                s += -1 * fE;
                s += -1 * fN;
                s += 1 * fx[kW];
                s += 1 * fy[kS];
                o->f[k] = aTP * o->fbp[k] - bTP * o->f[k] + o->dtv * s;
            }
}

/* ---------------- phase 9: cell macros ---------------- */

/* Update_MacroVar: HCZ Ark/Lib/D2Q9HCZ.cpp:133-182 (calcMu :24-40), AC-p
 * Ark/Lib/D2Q9AC.cpp:146-231 (calcMu :34-56), PP Ark/Lib/D2Q9PP.cpp:89-109 with
 * Update_PseudoPsi (Ark/Cpp/PseudoP.cpp:15-22). */
static void phase9_macros(ora *o)
{
    size_t N = o->N;
    const double rt = o->c.rt, hdt = o->hdt;
    double *m = o->m;
    for (size_t n = 0; n < N; ++n) {
        double s = 0, sx = 0, sy = 0;
        if (o->pp) {
            for (int q = 0; q < o->Q; ++q) s += o->f[q * N + n];
            for (int q = 0; q < o->Q; ++q) sx += o->f[q * N + n] * o->xi[0][q];
            for (int q = 0; q < o->Q; ++q) sy += o->f[q * N + n] * o->xi[1][q];
            double rho = s;
            double U = sx / rho, V = sy / rho;
            U += hdt * m[MPD_M_FX * N + n] / rho;
            V += hdt * m[MPD_M_FY * N + n] / rho;
            m[MPD_M_RHO * N + n] = rho;
            m[MPD_M_U * N + n] = U;
            m[MPD_M_V * N + n] = V;
            double p = o->c.pp_k * carnahan_starling(o, rho);
            m[MPD_M_P * N + n] = p;
            calc_psi(p, rho, rt, &m[MPD_M_SCG * N + n], &m[MPD_M_PSI * N + n]);
            continue;
        }
        double sh = 0;
        for (int q = 0; q < Q9; ++q) sh += o->h[q * N + n];
        double phi = sh;
        double rho = o->aphi * phi + o->bphi;
        for (int q = 0; q < Q9; ++q) s += o->f[q * N + n];
        double U, V, p, mu;
        double rx = m[MPD_M_RHOX * N + n], ry = m[MPD_M_RHOY * N + n];
        if (o->c.model == MPD_MODEL_AC_HCZ) {
            for (int q = 0; q < Q9; ++q) sx += o->f[q * N + n] * o->xi[0][q];
            for (int q = 0; q < Q9; ++q) sy += o->f[q * N + n] * o->xi[1][q];
            U = sx / (rho * rt);
            V = sy / (rho * rt);
            U += hdt * m[MPD_M_FX * N + n] / (rho * rt);
            V += hdt * m[MPD_M_FY * N + n] / (rho * rt);
            p = s + hdt * (rx * U + ry * V) * rt;
            mu = o->c.nu_v + (phi - o->c.phi_v) * (o->c.nu_l - o->c.nu_v);
            mu *= rho;
        } else {
            U = 0.0;
            V = 0.0;
            p = s - o->f[0 * N + n];
            p += (hdt * (rx * U + ry * V) - rho * W9[0] * (U * U + V * V) * o->lambda0);
            p *= o->kp;
            mu = rho * rt * o->c.tau0;
        }
        m[MPD_M_PHI * N + n] = phi;
        m[MPD_M_RHO * N + n] = rho;
        m[MPD_M_U * N + n] = U;
        m[MPD_M_V * N + n] = V;
        m[MPD_M_P * N + n] = p;
        m[MPD_M_MU * N + n] = mu;
        m[MPD_M_TAU * N + n] = mu / (rho * rt);
    }
}

/* ---------------- phase 10: Allen-Cahn sources ---------------- */

/* MacroSource (Ark/Cpp/SourceM.cpp:15-62): Grad_Phi_6points (GradScheme.cpp:347-376,
 * 9-point gradient GradSchemeBasic.h:153-180), SourceMomentum, and the
 * conservative AC source vector written over Phi_x/Phi_y. */
static void phase10_source(ora *o)
{
    int nx = o->nx, ny = o->ny;
    size_t N = o->N;
    double d2x = 2.0 * o->c.dx, minl = o->c.dx;
    const double rt = o->c.rt;
    const mpd_config *c = &o->c;
    /* gradients first, from phi of this step only (phi is not written here) */
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            size_t n = (size_t)j * nx + i;
            double C0 = M(o, MPD_M_PHI, i + 1, j), C1 = M(o, MPD_M_PHI, i, j + 1);
            double C2 = M(o, MPD_M_PHI, i - 1, j), C3 = M(o, MPD_M_PHI, i, j - 1);
            double D0 = M(o, MPD_M_PHI, i + 1, j + 1), D1 = M(o, MPD_M_PHI, i - 1, j + 1);
            double D2 = M(o, MPD_M_PHI, i - 1, j - 1), D3 = M(o, MPD_M_PHI, i + 1, j - 1);
            double phi = o->m[MPD_M_PHI * N + n];
            double gx = (4 * (C0 - C2) + 1 * (D0 - D1 + D3 - D2)) / (d2x * 6);
            double gy = (4 * (C1 - C3) + 1 * (D0 - D3 + D1 - D2)) / (d2x * 6);
            double lap = (4 * (C0 + C1 + C2 + C3) + (D0 + D1 + D2 + D3) - 20 * phi) /
                         (6 * minl * minl);
            set_zero(&gx);
            set_zero(&gy);
            /* SourceMomentum: chemical potential, F = mu_c grad(phi), grad(rho) */
            double mu_c = 2 * o->beta * (phi - c->phi_l) * (phi - c->phi_v) *
                              (2 * phi - c->phi_l - c->phi_v) -
                          o->kappa * lap;
            double rx = o->aphi * gx, ry = o->aphi * gy;
            double Fx = mu_c * gx, Fy = mu_c * gy;
            if (c->gx != 0.0) Fx -= c->gx * (o->m[MPD_M_RHO * N + n] - c->rho_ref);
            if (c->gy != 0.0) Fy -= c->gy * (o->m[MPD_M_RHO * N + n] - c->rho_ref);
            double U = o->m[MPD_M_U * N + n], V = o->m[MPD_M_V * N + n];
            double dpu = phi * U - o->m[MPD_M_PPU * N + n];
            double dpv = phi * V - o->m[MPD_M_PPV * N + n];
            double L = sqrt(gx * gx + gy * gy);
            double sx, sy;
            if (L != 0) {
                double tmp = -4 * (phi - c->phi_l) * (phi - c->phi_v) / (c->width * (c->phi_l - c->phi_v));
                sx = gx * tmp / L + dpu / (rt * c->dt);
                sy = gy * tmp / L + dpv / (rt * c->dt);
            } else {
                sx = dpu / (rt * c->dt);
                sy = dpv / (rt * c->dt);
            }
            o->m[MPD_M_LAP * N + n] = lap;
            o->m[MPD_M_RHOX * N + n] = rx;
            o->m[MPD_M_RHOY * N + n] = ry;
            o->m[MPD_M_FX * N + n] = Fx;
            o->m[MPD_M_FY * N + n] = Fy;
            o->m[MPD_M_PHIX * N + n] = sx;
            o->m[MPD_M_PHIY * N + n] = sy;
            o->m[MPD_M_PPU * N + n] = phi * U;
            o->m[MPD_M_PPV * N + n] = phi * V;
        }
}

/* ---------------- phase 11: pseudopotential cell force ---------------- */

/* Update_PseudoForce(Cell_2D&), Ark/Cpp/PseudoP.cpp:31-38 with the 9-point
 * gradient of Psi (Ark/Cpp/GradScheme.cpp:234-241). */
static void phase11_ppforce(ora *o)
{
    int nx = o->nx, ny = o->ny;
    size_t N = o->N;
    double d2x = 2.0 * o->c.dx;
    double *fxn = malloc(N * sizeof(double)), *fyn = malloc(N * sizeof(double));
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            size_t n = (size_t)j * nx + i;
            double C0 = M(o, MPD_M_PSI, i + 1, j), C1 = M(o, MPD_M_PSI, i, j + 1);
            double C2 = M(o, MPD_M_PSI, i - 1, j), C3 = M(o, MPD_M_PSI, i, j - 1);
            double D0 = M(o, MPD_M_PSI, i + 1, j + 1), D1 = M(o, MPD_M_PSI, i - 1, j + 1);
            double D2 = M(o, MPD_M_PSI, i - 1, j - 1), D3 = M(o, MPD_M_PSI, i + 1, j - 1);
            double gx = (4 * (C0 - C2) + 1 * (D0 - D1 + D3 - D2)) / (d2x * 6);
            double gy = (4 * (C1 - C3) + 1 * (D0 - D3 + D1 - D2)) / (d2x * 6);
            double sp = -o->m[MPD_M_SCG * N + n] * o->m[MPD_M_PSI * N + n];
            gx *= sp;
            gy *= sp;
            fxn[n] = gx;
            fyn[n] = gy;
        }
    memcpy(o->m + MPD_M_FX * N, fxn, N * sizeof(double));
    memcpy(o->m + MPD_M_FY * N, fyn, N * sizeof(double));
    free(fxn);
    free(fyn);
}

/* ---------------- one time step ---------------- */

int ora_step(ora *o, int nsteps)
{
    for (int s = 0; s < nsteps; ++s) {
        for (size_t n = 0; n < (size_t)o->N; ++n) phase1_cell(o, n);
        if (o->walls) phase2_walls(o);
        if (o->c.flux == MPD_FLUX_UW) {
            if (!o->pp) phase3_grad(o, o->hbp, o->ghbp, o->hbpx, o->hbpy);
            phase3_grad(o, o->fbp, o->gfbp, o->fbpx, o->fbpy);
        }
        phase4_faces(o);
        phase8_update(o);
        phase9_macros(o);
        if (o->pp) phase11_ppforce(o);
        else phase10_source(o);
    }
    return 0;
}

/* Wall-face net flux (fluxCheck, Ark/Cpp/Boundary.cpp:185-222): the largest
 * |sum_q h(x_f) xi.n dS| and |sum_q f(...)| over wall faces of the last step. */
double ora_wall_flux_residual(const ora *o)
{
    if (!o->walls) return 0.0;
    double worst = 0.0;
    size_t plane = (size_t)o->nyf * o->nx;
    const double *hy = o->hfl + (size_t)Q9 * o->N, *fy = o->ffl + (size_t)Q9 * o->N;
    for (int side = 0; side < 2; ++side) {
        int j = side == 0 ? 0 : o->ny;
        for (int i = 0; i < o->nx; ++i) {
            double sh = 0, sf = 0;
            for (int q = 0; q < Q9; ++q) {
                sh += hy[(size_t)q * plane + (size_t)j * o->nx + i];
                sf += fy[(size_t)q * plane + (size_t)j * o->nx + i];
            }
            if (fabs(sh) > worst) worst = fabs(sh);
            if (fabs(sf) > worst) worst = fabs(sf);
        }
    }
    return worst;
}

/* Phases 9-11 only (cell macros and sources from the current populations):
 * lets a test feed the reference's post-update populations and compare the
 * macros the oracle derives from them with the reference's. */
int ora_cell_phases(ora *o)
{
    phase9_macros(o);
    if (o->pp) phase11_ppforce(o);
    else phase10_source(o);
    return 0;
}
