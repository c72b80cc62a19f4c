// D2V16 discrete-velocity set (Ark/Lib/D2V16.h:14-20, Ark/Lib/D2V16.cpp:11-75):
// c = MaSpan = 1.2, eta = (3 + nK) c with nK = 0; four shells of four
// velocities: axis c, diagonal c, axis 2c, diagonal 2c.  The equilibrium is
// f_eq_j = sum_k A[j][k] M16_k (Update_phi_Eq, D2V16.cpp:157-166), A being the
// inverse of the 16x16 moment matrix of the basis
//   1, xi_x, xi_y, E, xi_x xi_y, xi_x^2, xi_y^2, xi_x E, xi_y E, xi_x^3, xi_y^3,
//   xi_x^2 xi_y, xi_x xi_y^2, xi_x xi_y E, xi_x^2 E, xi_y^2 E
// with E = |xi|^2 + eta^2 on the first shell and |xi|^2 on the others
// (xi2Eta2, D2V16.cpp:68-74).  Rows of A are built here from the closed-form
// coefficients of D2V16.cpp:17-35 and the sign pattern of each shell.
// Host and device code (constexpr, no HIP types).
#pragma once

namespace mpd {
namespace d2v16 {

constexpr int Q = 16;
constexpr double C = 1.2;                 // MaSpan
constexpr double ETA = (3.0 + 0.0) * C;   // (3 + nK) MaSpan
constexpr double DEG = 3.0 + 0.0;         // degrees of freedom, 3 + nK (D2V16.cpp:5)
constexpr double MAXU = 2 * C * 1.4142135623730951;  // 2 MaSpan sqrt(2)

// shell s = j / 4; direction within the shell (unit signs)
constexpr int shell(int j) { return j / 4; }
constexpr int ux(int j)
{
    return shell(j) % 2 == 0 ? ((j % 4) == 0 ? 1 : (j % 4) == 2 ? -1 : 0)
                             : ((j % 4) == 0 || (j % 4) == 3 ? 1 : -1);
}
constexpr int uy(int j)
{
    return shell(j) % 2 == 0 ? ((j % 4) == 1 ? 1 : (j % 4) == 3 ? -1 : 0)
                             : ((j % 4) <= 1 ? 1 : -1);
}
constexpr double speed(int j) { return shell(j) < 2 ? C : 2 * C; }
constexpr double xi_x(int j) { return speed(j) * ux(j); }
constexpr double xi_y(int j) { return speed(j) * uy(j); }
// E_j = |xi_j|^2 (+ eta^2 on the first shell)
constexpr double energy(int j)
{
    return shell(j) == 0 ? C * C + ETA * ETA : shell(j) == 1 ? 2 * C * C : shell(j) == 2 ? 4 * C * C : 8 * C * C;
}

struct Coef {
    double A, B, D, E, F, G, H, I, J, K, L, N, P, Qc, R, S, T, U, V;
};
constexpr Coef coef()
{
    const double c = C, e = ETA;
    const double c2 = c * c, e2 = e * e, c2e2 = c2 * e2;
    Coef k{};
    k.A = 1.0 / (4.0 * e2);
    k.B = (-21.0 * c2 - e2) / (48.0 * c2e2);
    k.D = (7.0 * c2 + e2) / (32.0 * c2e2);
    k.E = (-3.0 * c2 - e2) / (96.0 * c2e2);
    k.F = 1.0 / (3.0 * c2);
    k.G = (3.0 * c2 - 5.0 * e2) / (4.0 * e2 * (e2 - 3.0 * c2));
    k.H = 3.0 * (c2 + e2) / (4.0 * e2 * (e2 - 3.0 * c2));
    k.I = (21.0 * c2 - 11.0 * e2) / (48.0 * c2e2);
    k.J = (21.0 * c2 * c2 - 32.0 * c2e2 + 11.0 * e2 * e2) / (32.0 * c2e2 * (e2 - 3.0 * c2));
    k.K = (21.0 * c2 * c2 - 36.0 * c2e2 + 7.0 * e2 * e2) / (32.0 * c2e2 * (e2 - 3.0 * c2));
    k.L = (3.0 * c2 - 5.0 * e2) / (96.0 * c2e2);
    k.N = 1.0 / (2.0 * c * e2);
    k.P = (3.0 * c2 - e2) / (12.0 * c2e2 * c);
    k.Qc = (c2 + e2) / (16.0 * c2e2 * c);
    k.R = (c2 - e2) / (16.0 * c2e2 * c);
    k.S = 1.0 / (24.0 * c2 * c2);
    k.T = 1.0 / (4.0 * c2 * (e2 - 3.0 * c2));
    k.U = (c2 - e2) / (32.0 * c2 * c2 * (e2 - 3.0 * c2));
    k.V = (5.0 * c2 - e2) / (32.0 * c2 * c2 * (e2 - 3.0 * c2));
    return k;
}

// entry A[j][m] of the inverse moment matrix
constexpr double inv(int j, int m)
{
    const Coef k = coef();
    const double x = ux(j), y = uy(j), xx = x * x, yy = y * y, c = C;
    switch (shell(j)) {
    case 0: {  // (+-c, 0), (0, +-c)
        const double r[Q] = {0, 0, 0, k.A, 0, xx * k.G + yy * k.H, xx * k.H + yy * k.G, x * k.N, y * k.N,
                             -x * k.N, -y * k.N, -y * k.N, -x * k.N, 0, (xx - yy) * k.T, (yy - xx) * k.T};
        return r[m];
    }
    case 1: {  // (+-c, +-c)
        const double r[Q] = {2.0 / 3.0, x * (1.0 / (3.0 * c)), y * (1.0 / (3.0 * c)), k.B, x * y * k.F, k.I, k.I,
                             -x * (k.N / 2), -y * (k.N / 2), x * k.P, y * k.P, y * (k.N / 2), x * (k.N / 2),
                             -x * y * k.S, k.S / 2, k.S / 2};
        return r[m];
    }
    case 2: {  // (+-2c, 0), (0, +-2c)
        const double r[Q] = {-0.5, 0, 0, k.D, 0, xx * k.J + yy * k.K, xx * k.K + yy * k.J, -x * (k.N / 8),
                             -y * (k.N / 8), x * k.Qc, y * k.Qc, y * k.R, x * k.R, 0, xx * k.U + yy * k.V,
                             xx * k.V + yy * k.U};
        return r[m];
    }
    default: {  // (+-2c, +-2c)
        const double r[Q] = {1.0 / 12, -x * (1.0 / (24 * c)), -y * (1.0 / (24 * c)), k.E, -x * y * (k.F / 16),
                             k.L, k.L, x * (k.N / 16), y * (k.N / 16), -x * (k.P / 8), -y * (k.P / 8),
                             -y * (k.R / 2), -x * (k.R / 2), x * y * (k.S / 4), k.S / 4, k.S / 4};
        return r[m];
    }
    }
}

// the whole table, for code that indexes it (compile-time evaluated)
struct Table {
    double a[Q][Q];
    double x[Q], y[Q], e[Q];
};
constexpr Table make_table()
{
    Table t{};
    for (int j = 0; j < Q; ++j) {
        for (int m = 0; m < Q; ++m) t.a[j][m] = inv(j, m);
        t.x[j] = xi_x(j);
        t.y[j] = xi_y(j);
        t.e[j] = energy(j);
    }
    return t;
}

// Moments of the equilibrium (Macro16, D2V16.cpp:125-145) at temperature RT
// (gas constant times T) and pressure p.
template <typename T>
constexpr void macro16(T *M, T rho, T u, T v, T rt, T p)
{
    const T u2 = u * u, v2 = v * v;
    const T deg = (T)DEG, degp2 = (T)(DEG + 2.0);
    M[0] = rho;
    M[1] = rho * u;
    M[2] = rho * v;
    M[3] = rho * (deg * rt + u2 + v2);
    M[4] = M[1] * v;
    M[5] = rho * u2 + p;
    M[6] = rho * v2 + p;
    M[7] = M[1] * (degp2 * rt + u2 + v2);
    M[8] = M[2] * (degp2 * rt + u2 + v2);
    M[9] = M[1] * ((T)3.0 * rt + u2);
    M[10] = M[2] * ((T)3.0 * rt + v2);
    M[11] = M[2] * (rt + u2);
    M[12] = M[1] * (rt + v2);
    M[13] = ((deg + (T)4.0) * rt + u2 + v2) * M[4];
    M[14] = degp2 * rho * rt * rt + p * ((deg + (T)5.0) * u2 + v2) + rho * u2 * (u2 + v2);
    M[15] = degp2 * rho * rt * rt + p * ((deg + (T)5.0) * v2 + u2) + rho * v2 * (u2 + v2);
}

}  // namespace d2v16
}  // namespace mpd
