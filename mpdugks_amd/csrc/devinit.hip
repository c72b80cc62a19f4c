// Start states built on the GPU, straight into a handle's FP64 state arrays:
// no host arrays and no upload (a 1024x1024 FP64 state is ~300 MB over PCIe).
// They follow the host initialisers of hostinit.cpp (AC_Drop,
// Ark/Cpp/Preprocess.cpp:705-760 with AnalyticalPhiAC_Drop :402-419; the
// interface layer; SCMP_Drop and SCMP_FlatInterface, :291-401; the noisy
// pseudopotential start), with the device's tanh
// and sqrt: the states agree with the host ones to a few units in the last
// place, not bit for bit.  Compiled without contraction.
#include "d2v16.h"
#include "device_state.h"
#include "physics.h"

namespace mpd {

namespace {

constexpr d2v16::Table kV16i = d2v16::make_table();

// equilibria at rest (Update_phi_Eq with zero sources) of cell n, written as the tilde state
__device__ inline void write_rest_state(const DevState &S, const Consts &C, size_t n, double phi, double rho, double p,
                                        double mu, double tau)
{
    const size_t N = S.N;
    double *m = S.m;
    for (int k = 0; k < MPD_M_COUNT; ++k) m[(size_t)k * N + n] = 0.0;
    m[MPD_M_PHI * N + n] = phi;
    m[MPD_M_RHO * N + n] = rho;
    m[MPD_M_P * N + n] = p;
    m[MPD_M_MU * N + n] = mu;
    m[MPD_M_TAU * N + n] = tau;
    S.rho1k[n] = 0.0;  // Rho_1k starts at 0 (Ark/Header/Macroscopic.h:44)
    if (C.model == MPD_MODEL_PP_D2V16) {
        double M[d2v16::Q];
        d2v16::macro16<double>(M, rho, 0.0, 0.0, C.rt, rho * C.rt);
        for (int j = 0; j < d2v16::Q; ++j) {
            double s = 0;
            for (int k = 0; k < d2v16::Q; ++k) s += kV16i.a[j][k] * M[k];
            S.f[(size_t)j * N + n] = s;
        }
        return;
    }
    const double rt = C.rt;
    for (int q = 0; q < Q9; ++q) {
        const double w = C.w[q];
        const double u1 = 0.0, G = u1 + 0.5 * u1 * u1 - 0.0 * C.lambda0;
        if (C.model == MPD_MODEL_PP_D2Q9) {
            S.f[(size_t)q * N + n] = w * rho * (1.0 + u1 + 0.5 * u1 * u1 - 0.0 * C.lambda0);
            continue;
        }
        S.h[(size_t)q * N + n] = w * (1.0 + u1) * phi;
        S.f[(size_t)q * N + n] = C.model == MPD_MODEL_AC_HCZ ? w * (p + rho * G * rt)
                                                             : w * (p / rt + rho * G) - (double)(q == 0) * p / rt;
    }
}

__global__ void __launch_bounds__(256) k_init_drops(DevState S, Consts C, DropSet D, int sign, double p0, double mu0)
{
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (size_t)S.N) return;
    const int i = (int)(n % C.nx), j = (int)(n / C.nx);
    const double x = (i + 0.5) * C.dx, y = (j + 0.5) * C.dx;
    double phi = 0.0;
    for (int d = 0; d < D.n; ++d) {
        double t = tanh((sqrt((x - D.cx[d]) * (x - D.cx[d]) + (y - D.cy[d]) * (y - D.cy[d])) - D.r[d]) * 2 / D.width);
        if (sign < 0) t = -t;
        const double p = 0.5 * (C.phi_l + C.phi_v) + 0.5 * (C.phi_l - C.phi_v) * t;
        // droplets (sign -1) join as the larger phi, bubbles as the smaller
        phi = d == 0 ? p : (sign < 0 ? fmax(phi, p) : fmin(phi, p));
    }
    const double rho = C.aphi * phi + C.bphi;
    write_rest_state(S, C, n, phi, rho, p0, mu0, mu0 / (rho * C.rt));
}

__global__ void __launch_bounds__(256) k_init_interface(DevState S, Consts C, const double *y0, int sign, double p0,
                                                        double width)
{
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (size_t)S.N) return;
    const int i = (int)(n % C.nx), j = (int)(n / C.nx);
    const double y = (j + 0.5) * C.dx;
    double t = tanh(2.0 * (y - y0[i]) / width);
    if (sign < 0) t = -t;
    const double phi = 0.5 * (C.phi_l + C.phi_v) + 0.5 * (C.phi_l - C.phi_v) * t;
    const double rho = C.aphi * phi + C.bphi;
    const double mu = (C.nu_v + (phi - C.phi_v) * (C.nu_l - C.nu_v)) * rho;
    write_rest_state(S, C, n, phi, rho, p0, mu, mu / (rho * C.rt));
}

// splitmix64 (Steele, Lea, Flood 2014) of (seed, cell): a counter-based stream,
// so every cell draws independently of the launch shape
__device__ inline uint64_t splitmix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__global__ void __launch_bounds__(256) k_init_noise(DevState S, Consts C, double rho0, double amp, uint64_t seed,
                                                    double mu0)
{
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (size_t)S.N) return;
    const double u = (double)(splitmix64(seed ^ splitmix64((uint64_t)n)) >> 11) * 0x1.0p-53;  // [0, 1)
    const double rho = rho0 * (1.0 + (-amp + 2.0 * amp * u));
    write_rest_state(S, C, n, 0.0, rho, rho * C.rt, mu0, C.pp_tau);
}

// SCMP_Drop (Ark/Cpp/Preprocess.cpp:337-401), the expression of mpd_init_pp_drop
__global__ void __launch_bounds__(256) k_init_pp_drop(DevState S, Consts C, double cx, double cy, double radius,
                                                      double rho_l, double rho_v, double width, double mu0)
{
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (size_t)S.N) return;
    const int i = (int)(n % C.nx), j = (int)(n / C.nx);
    const double x = (i + 0.5) * C.dx, y = (j + 0.5) * C.dx;
    const double rho = 0.5 * (rho_l + rho_v) -
                       0.5 * (-rho_l + rho_v) * tanh(2 * (sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy)) - radius) / width);
    write_rest_state(S, C, n, 0.0, rho, rho * C.rt, mu0, C.pp_tau);
}

// SCMP_FlatInterface (Ark/Cpp/Preprocess.cpp:291-336), the expression of mpd_init_pp_flat
__global__ void __launch_bounds__(256) k_init_pp_flat(DevState S, Consts C, double y_lo, double y_hi, double slope,
                                                      double rho_l, double rho_v, double mu0)
{
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (size_t)S.N) return;
    const double y = ((int)(n / C.nx) + 0.5) * C.dx;
    const double rho = rho_v + 0.5 * (rho_l - rho_v) * (tanh(slope * (y - y_lo)) - tanh(slope * (y - y_hi)));
    write_rest_state(S, C, n, 0.0, rho, rho * C.rt, mu0, C.pp_tau);
}

__global__ void k_zero(double *p, size_t n)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0.0;
}

inline unsigned blocks(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

void devinit_drops(const DevState &S, const Consts &C, const DropSet &D, int sign, double p0, double mu0,
                   hipStream_t st)
{
    hipLaunchKernelGGL(k_init_drops, dim3(blocks(S.N)), dim3(256), 0, st, S, C, D, sign, p0, mu0);
}

void devinit_interface(const DevState &S, const Consts &C, const double *d_y0, int sign, double p0, double width,
                       hipStream_t st)
{
    hipLaunchKernelGGL(k_init_interface, dim3(blocks(S.N)), dim3(256), 0, st, S, C, d_y0, sign, p0, width);
}

// pseudopotential face forces, and the face Psi the reference face order keeps
// from the previous step, start at zero (as mpd_upload_state leaves them)
static void zero_pp_faces(const DevState &S, const Consts &C, hipStream_t st)
{
    double *planes[] = {S.ffx, S.ffy, S.fpsi_old};
    for (double *p : planes)
        if (p) hipLaunchKernelGGL(k_zero, dim3(blocks(2 * S.N)), dim3(256), 0, st, p, 2 * (size_t)S.N);
    // the second copies of the periodic faces (UW, reference face order)
    const size_t nb = (size_t)C.nx + C.ny;
    double *copies[] = {S.cbffx, S.cbffy, S.cbpsi_old};
    for (double *p : copies)
        if (p) hipLaunchKernelGGL(k_zero, dim3(blocks(nb)), dim3(256), 0, st, p, nb);
}

void devinit_noise(const DevState &S, const Consts &C, double rho0, double amp, uint64_t seed, double mu0,
                   hipStream_t st)
{
    hipLaunchKernelGGL(k_init_noise, dim3(blocks(S.N)), dim3(256), 0, st, S, C, rho0, amp, seed, mu0);
    zero_pp_faces(S, C, st);
}

void devinit_pp_drop(const DevState &S, const Consts &C, double cx, double cy, double radius, double rho_l,
                     double rho_v, double width, double mu0, hipStream_t st)
{
    hipLaunchKernelGGL(k_init_pp_drop, dim3(blocks(S.N)), dim3(256), 0, st, S, C, cx, cy, radius, rho_l, rho_v,
                       width, mu0);
    zero_pp_faces(S, C, st);
}

void devinit_pp_flat(const DevState &S, const Consts &C, double y_lo, double y_hi, double slope, double rho_l,
                     double rho_v, double mu0, hipStream_t st)
{
    hipLaunchKernelGGL(k_init_pp_flat, dim3(blocks(S.N)), dim3(256), 0, st, S, C, y_lo, y_hi, slope, rho_l, rho_v,
                       mu0);
    zero_pp_faces(S, C, st);
}

}  // namespace mpd
