// C ABI of the DUGKS step library (include/mpdugks.h): handle life cycle,
// state upload/download, stepping and global reductions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/mpdugks.h"
#include "device_state.h"
#include "d2v16.h"
#include "fused.h"
#include "last_error.h"
#include "physics.h"

using namespace mpd;

namespace {

thread_local std::string g_err;

int vfail(int code, const char *fmt, va_list ap)
{
    char buf[512];
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    g_err = buf;
    return code;
}

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vfail(code, fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) return fail(MPD_ERR_DEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

}  // namespace

int mpd::set_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vfail(code, fmt, ap);
    va_end(ap);
    return code;
}

struct mpd_batch;

struct mpd_handle {
    mpd_config cfg;
    Consts C;
    DevState S;
    FusedState F;
    hipStream_t stream = nullptr;
    int device = 0;
    int cus = 0;                 // compute units of the device
    int tile = MPD_TILE_AUTO;    // FP32-bulk Allen-Cahn tile shape (mpd_tile)
    int path = MPD_PATH_AUTO;
    int effective_path = MPD_PATH_PHASES;
    bool uploaded = false;
    bool fused_primed = false;   // fused state derived from the explicit macros
    int prime_stage = 0;         // explicit steps taken towards priming
    double tau_u = 0.0;          // pseudopotential: the relaxation time of every cell, when the state has one
    std::vector<void *> allocs;
    int64_t bytes = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double acc_ms = 0.0;
    int64_t launches = 0;
    bool per_launch = false;
    std::vector<hipEvent_t> evs;  // per-launch timing pairs
    // mpd_step brackets its launches with an event pair and returns; the pairs
    // are read (and recycled) when the caller synchronises or asks for timing
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    std::vector<hipEvent_t> ev_free;
    double *d_part = nullptr;
    int red_blocks = 0;
    mpd_batch *batch = nullptr;  // set while the handle is a member of a batch
};

struct mpd_batch {
    hipStream_t stream = nullptr;
    std::vector<mpd_handle *> members;
    Batch32Plan *plan = nullptr;
    bool dirty = true;            // tables / graph need rebuilding
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double acc_ms = 0.0;
    int64_t steps = 0;
    int tile_req = MPD_TILE_AUTO;   // tile shape: tuned, or forced (mpd_tile)
    int tile = MPD_TILE_16X15;      // shape of the current tables
    bool tune_pending = false;      // members changed since the last tuning
    double tune_ms[3] = {0.0, 0.0, 0.0};  // last tuning: ms per step of each shape
};

namespace {

int dalloc(mpd_handle *h, double **p, size_t n)
{
    *p = nullptr;
    if (n == 0) return 0;
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, n * sizeof(double));
    if (e != hipSuccess) return fail(MPD_ERR_NOMEM, "hipMalloc(%zu doubles): %s", n, hipGetErrorString(e));
    h->allocs.push_back(q);  // freed by mpd_destroy, also if the clearing below fails
    h->bytes += (int64_t)(n * sizeof(double));
    e = hipMemsetAsync(q, 0, n * sizeof(double), h->stream);
    if (e != hipSuccess) return fail(MPD_ERR_DEVICE, "hipMemsetAsync: %s", hipGetErrorString(e));
    *p = (double *)q;
    return 0;
}

void make_consts(const mpd_config &c, Consts &C)
{
    std::memset(&C, 0, sizeof(C));
    C.nx = c.nx;
    C.ny = c.ny;
    C.model = c.model;
    C.flux = c.flux;
    C.walls = c.bc_y == MPD_BC_WALL_BB || c.bc_y == MPD_BC_WALL_NEE;
    C.nee = c.bc_y == MPD_BC_WALL_NEE;
    C.nyf = C.walls ? c.ny + 1 : c.ny;
    C.dx = c.dx;
    C.dt = c.dt;
    C.hdt = 0.5 * c.dt;
    C.rt = c.rt;
    C.lambda0 = 1 / (2.0 * c.rt);
    C.aphi = (c.rho_l - c.rho_v) / (c.phi_l - c.phi_v);
    C.bphi = c.rho_v - C.aphi * c.phi_v;
    C.beta = 12.0 * c.sigma / c.width;
    C.kappa = 3.0 * c.sigma * c.width / 2;
    C.kp = c.rt * 9 / 5.0;
    C.taum = c.mobility / c.rt;
    C.tau0 = c.tau0;
    C.phi_l = c.phi_l;
    C.phi_v = c.phi_v;
    C.nu_l = c.nu_l;
    C.nu_v = c.nu_v;
    C.gx = c.gx;
    C.gy = c.gy;
    C.rho_ref = c.rho_ref;
    C.pp_k = c.pp_k;
    C.t_cs = c.pp_tr * 0.0943;
    C.pp_tau = c.pp_tau;
    C.dtv = c.dt / (c.dx * c.dx);
    C.d2x6 = (2.0 * c.dx) * 6;
    C.lapden = 6 * c.dx * c.dx;
    C.rtdt = c.rt * c.dt;
    C.wdphi = c.width * (c.phi_l - c.phi_v);
    C.twodx = 2 * c.dx;
    C.rdx = 1.0 / c.dx;
    C.rtwodx = 1.0 / C.twodx;
    C.ah_h = 2.0 * C.taum / (2.0 * C.taum + C.hdt);
    C.bh_h = 1.0 - C.ah_h;
    C.ch_h = C.taum * C.bh_h;
    C.abp_h = (2.0 * C.taum - C.hdt) / (2.0 * C.taum + c.dt);
    C.bbp_h = 1.0 - C.abp_h;
    C.cbp_h = C.taum * C.bbp_h;
    C.ah_pp = 2.0 * c.pp_tau / (2.0 * c.pp_tau + C.hdt);
    C.bh_pp = 1.0 - C.ah_pp;
    C.ch_pp = c.pp_tau * C.bh_pp;
    double maspan = std::sqrt(3.0 * c.rt);
    const double w9[Q9] = {4.0 / 9.0, 1.0 / 9.0, 1.0 / 36.0, 1.0 / 9.0, 1.0 / 36.0,
                           1.0 / 9.0, 1.0 / 36.0, 1.0 / 9.0, 1.0 / 36.0};
    for (int q = 0; q < Q9; ++q) {
        C.xi[0][q] = maspan * ex9(q);
        C.xi[1][q] = maspan * ey9(q);
        C.w[q] = w9[q];
    }
    // unit lattice: dx a normal power of two whose reciprocal (and half of it)
    // is normal too, and the lattice velocities exactly 0 and +-1
    int e = 0;
    C.unit = std::frexp(c.dx, &e) == 0.5 && e > -1000 && e < 1000;
    for (int q = 0; q < Q9; ++q) C.unit = C.unit && C.xi[0][q] == ex9(q) && C.xi[1][q] == ey9(q);
}

// ---------------- reductions ----------------

constexpr int RED_THREADS = 256;
constexpr int RED_FIELDS = 9;  // rho, rho u, rho v, phi, KE, max|u|, dRho^2, rho^2, nonfinite

__device__ inline void neumaier(double &s, double &c, double x)
{
    double t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
}

__global__ void __launch_bounds__(RED_THREADS) k_reduce(const double *m, double *rho1k, int64_t N,
                                                        int update, double *part)
{
    double s[RED_FIELDS - 1] = {0}, c[RED_FIELDS - 1] = {0};
    double vmax = 0.0, bad = 0.0;
    for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < N;
         n += (int64_t)gridDim.x * blockDim.x) {
        double rho = m[MPD_M_RHO * N + n], U = m[MPD_M_U * N + n], V = m[MPD_M_V * N + n];
        double phi = m[MPD_M_PHI * N + n];
        bool ok = isfinite(rho) && isfinite(U) && isfinite(V) && isfinite(phi) && rho > 0.0;
        if (!ok) { bad += 1.0; continue; }
        double d = rho - rho1k[n];
        neumaier(s[0], c[0], rho);
        neumaier(s[1], c[1], rho * U);
        neumaier(s[2], c[2], rho * V);
        neumaier(s[3], c[3], phi);
        neumaier(s[4], c[4], 0.5 * rho * (U * U + V * V));
        neumaier(s[5], c[5], d * d);
        neumaier(s[6], c[6], rho * rho);
        vmax = fmax(vmax, sqrt(U * U + V * V));
        if (update) rho1k[n] = rho;
    }
    // across the 64 lanes of each wavefront: butterfly shuffles, each level
    // adding the partner's (sum, compensation) pair with an exact two-sum
    for (int off = 32; off > 0; off >>= 1) {
        for (int k = 0; k < 7; ++k) {
            const double so = __shfl_xor(s[k], off, 64), co = __shfl_xor(c[k], off, 64);
            neumaier(s[k], c[k], so);
            c[k] += co;
        }
        vmax = fmax(vmax, __shfl_xor(vmax, off, 64));
        bad += __shfl_xor(bad, off, 64);
    }
    // across the wavefronts of the workgroup
    constexpr int NW = RED_THREADS / 64;
    __shared__ double sh[RED_FIELDS][NW];
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    if (lane == 0) {
        for (int k = 0; k < 7; ++k) sh[k][wave] = s[k] + c[k];
        sh[7][wave] = vmax;
        sh[8][wave] = bad;
    }
    __syncthreads();
    if (threadIdx.x < RED_FIELDS) {
        const int k = threadIdx.x;
        double t = sh[k][0], e = 0.0;
        for (int w = 1; w < NW; ++w) {
            if (k == 7) t = fmax(t, sh[k][w]);
            else neumaier(t, e, sh[k][w]);
        }
        part[(size_t)blockIdx.x * RED_FIELDS + k] = t + e;
    }
}

int num_q(const mpd_config &c) { return c.model == MPD_MODEL_PP_D2V16 ? d2v16::Q : Q9; }
bool is_pp(const mpd_config &c) { return c.model == MPD_MODEL_PP_D2Q9 || c.model == MPD_MODEL_PP_D2V16; }

int check_cfg(const mpd_config *c)
{
    if (!c) return fail(MPD_ERR_INVALID, "null config");
    if (c->nx < 4 || c->ny < 4) return fail(MPD_ERR_INVALID, "mesh %dx%d too small (need >= 4x4)", c->nx, c->ny);
    if ((int64_t)c->nx * c->ny > (int64_t)1 << 30) return fail(MPD_ERR_INVALID, "mesh too large");
    if (c->model < MPD_MODEL_AC_P || c->model > MPD_MODEL_PP_D2V16)
        return fail(MPD_ERR_INVALID, "unknown model %d", c->model);
    if (c->model == MPD_MODEL_PP_D2V16 && (c->flux != MPD_FLUX_CD || c->bc_y != MPD_BC_PERIODIC))
        return fail(MPD_ERR_INVALID, "D2V16 runs with CD faces on a periodic mesh");
    if (c->flux != MPD_FLUX_CD && c->flux != MPD_FLUX_UW) return fail(MPD_ERR_INVALID, "unknown flux %d", c->flux);
    if (c->bc_y != MPD_BC_PERIODIC && c->bc_y != MPD_BC_WALL_BB && c->bc_y != MPD_BC_WALL_NEE)
        return fail(MPD_ERR_INVALID, "unknown bc %d", c->bc_y);
    if (c->bc_y != MPD_BC_PERIODIC && c->model == MPD_MODEL_PP_D2Q9)
        return fail(MPD_ERR_INVALID, "walls are supported with the Allen-Cahn models only");
    if (c->precision != MPD_FP64 && c->precision != MPD_FP32_BULK)
        return fail(MPD_ERR_INVALID, "unknown precision %d", c->precision);
    if (c->precision == MPD_FP32_BULK && !fused_supported(*c))
        return fail(MPD_ERR_INVALID, "no FP32-bulk kernel for this model / flux / boundary (FP32-bulk: Allen-Cahn "
                                     "with CD or UW faces and periodic, bounce-back or NEE y; pseudopotential D2Q9 "
                                     "with CD or UW faces, D2V16 with CD faces, periodic)");
    if (c->precision == MPD_FP32_BULK && !is_pp(*c)) {
        // the FP32 step kernel addresses each state set through one 32-bit buffer
        // resource (all float planes) and the FP64 phase field through another
        const int64_t n = (int64_t)c->nx * c->ny, lim = (int64_t)1 << 32;
        if (4 * n * fused_nplanes(*c) >= lim || 8 * n >= lim)
            return fail(MPD_ERR_INVALID, "mesh %dx%d too large for the FP32-bulk Allen-Cahn kernel (%d float "
                                         "planes of 4 B per cell must span < 4 GiB: at most %lld cells)",
                        c->nx, c->ny, fused_nplanes(*c), (long long)((lim - 1) / (4 * fused_nplanes(*c))));
    }
    if (c->precision == MPD_FP32_BULK && is_pp(*c)) {
        // the FP32 pseudopotential step addresses a state set's float planes, its
        // FP64 density and the shared planes through 32-bit buffer resources
        const int64_t n = (int64_t)c->nx * c->ny, lim = (int64_t)1 << 32;
        const int np = std::max(pp32_set_planes(num_q(*c)), 7);
        if (4 * n * np >= lim || 8 * n >= lim)
            return fail(MPD_ERR_INVALID, "mesh %dx%d too large for the FP32-bulk pseudopotential kernel (%d float "
                                         "planes of 4 B per cell must span < 4 GiB: at most %lld cells)",
                        c->nx, c->ny, np, (long long)((lim - 1) / (4 * np)));
    }
    if (!(c->dx > 0) || !(c->dt > 0) || !(c->rt > 0)) return fail(MPD_ERR_INVALID, "dx, dt, rt must be > 0");
    if (c->phi_l == c->phi_v || !(c->width > 0)) return fail(MPD_ERR_INVALID, "phi_l == phi_v or width <= 0");
    return 0;
}

}  // namespace

extern "C" {

int mpd_abi_version(void) { return MPDUGKS_ABI_VERSION; }

const char *mpd_last_error(void) { return g_err.c_str(); }

int mpd_default_config(int32_t model, mpd_config *c)
{
    if (!c) return fail(MPD_ERR_INVALID, "null config");
    std::memset(c, 0, sizeof(*c));
    const double ch_length = 128.0, min_l = ch_length / 128;
    const double t0 = 1.0 / 3, r0 = 1.0, rt = r0 * t0;
    const double mu0 = 0.1, rho0 = 1.0, nu0 = mu0 / rho0, tau0 = nu0 / rt;
    const double maxu = std::sqrt(3.0 * rt), cfl = 0.5;
    const double cn = 5 / ch_length;
    c->nx = 128;
    c->ny = 128;
    c->model = model;
    c->flux = MPD_FLUX_CD;
    c->bc_y = MPD_BC_PERIODIC;
    c->precision = MPD_FP64;
    c->dx = min_l;
    c->dt = cfl > 0.0 ? cfl * min_l / maxu : 1.0e-4;
    c->rt = rt;
    c->phi_l = 1.0;
    c->phi_v = 0.0;
    c->rho_l = 10.0;
    c->rho_v = 1.0;
    c->nu_l = nu0;
    c->nu_v = nu0;
    c->mobility = 0.1;
    c->width = ch_length * cn;
    c->sigma = 1e-2;
    c->tau0 = tau0;
    c->pp_tr = 0.85;
    c->pp_k = 1.0;
    c->pp_tau = tau0;
    if (model == MPD_MODEL_PP_D2V16) c->dt = cfl * min_l / d2v16::MAXU;  // MaxU = 2 MaSpan sqrt(2), D2V16.h
    return 0;
}

int mpd_create(const mpd_config *cfg, mpd_handle **out)
{
    if (!out) return fail(MPD_ERR_INVALID, "null output pointer");
    *out = nullptr;
    if (int r = check_cfg(cfg)) return r;
    mpd_handle *h = new (std::nothrow) mpd_handle();
    if (!h) return fail(MPD_ERR_NOMEM, "host allocation failed");
    h->cfg = *cfg;
    make_consts(*cfg, h->C);
    int r = 0;
    hipError_t e = hipGetDevice(&h->device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&h->cus, hipDeviceAttributeMultiprocessorCount, h->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&h->ev0);
    if (e == hipSuccess) e = hipEventCreate(&h->ev1);
    if (e != hipSuccess) {
        r = fail(MPD_ERR_DEVICE, "HIP set-up failed: %s", hipGetErrorString(e));
        mpd_destroy(h);
        return r;
    }
    const size_t N = (size_t)cfg->nx * cfg->ny, QN = (size_t)num_q(*cfg) * N;
    const size_t yfn = (size_t)h->C.nyf * cfg->nx;
    DevState &S = h->S;
    S.N = (int64_t)N;
    bool pp = is_pp(*cfg), ac = !pp, uw = cfg->flux == MPD_FLUX_UW;
    if ((r = dalloc(h, &S.h, ac ? QN : 0)) || (r = dalloc(h, &S.f, QN)) ||
        (r = dalloc(h, &S.m, (size_t)MPD_M_COUNT * N)) || (r = dalloc(h, &S.ffx, pp ? 2 * N : 0)) ||
        (r = dalloc(h, &S.ffy, pp ? 2 * N : 0)) || (r = dalloc(h, &S.fpsi_old, pp ? 2 * N : 0)) ||
        (r = dalloc(h, &S.rho1k, N))) {
        mpd_destroy(h);
        return r;
    }
    // FP64 pseudopotential D2Q9 with UW faces: the second copies of the periodic
    // faces, carried for the reference face order (their previous Psi and force)
    const bool pp_uw = cfg->model == MPD_MODEL_PP_D2Q9 && uw && cfg->precision == MPD_FP64;
    const size_t nb = (size_t)cfg->nx + cfg->ny;
    if ((r = dalloc(h, &S.cbpsi_old, pp_uw ? nb : 0)) || (r = dalloc(h, &S.cbffx, pp_uw ? nb : 0)) ||
        (r = dalloc(h, &S.cbffy, pp_uw ? nb : 0))) {
        mpd_destroy(h);
        return r;
    }
    // phase-path scratch is allocated lazily in mpd_step (fused path does not need it)
    (void)QN; (void)yfn; (void)uw;
    h->red_blocks = 512;
    if ((r = dalloc(h, &h->d_part, (size_t)h->red_blocks * RED_FIELDS))) {
        mpd_destroy(h);
        return r;
    }
    *out = h;
    return 0;
}

int mpd_batch_remove(mpd_batch *b, mpd_handle *h);

int mpd_destroy(mpd_handle *h)
{
    if (!h) return 0;
    if (h->batch) mpd_batch_remove(h->batch, h);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (void *p : h->allocs) (void)hipFree(p);
    for (auto e : h->evs) (void)hipEventDestroy(e);
    for (auto &pr : h->pending) {
        (void)hipEventDestroy(pr.first);
        (void)hipEventDestroy(pr.second);
    }
    for (auto e : h->ev_free) (void)hipEventDestroy(e);
    fused_free(h->F);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return 0;
}

int mpd_set_path(mpd_handle *h, int32_t path)
{
    if (!h) return fail(MPD_ERR_INVALID, "null handle");
    if (h->batch) return fail(MPD_ERR_STATE, "handle is a member of a batch");
    if (path < MPD_PATH_AUTO || path > MPD_PATH_FUSED) return fail(MPD_ERR_INVALID, "unknown path %d", path);
    if (path == MPD_PATH_FUSED && !fused_supported(h->cfg))
        return fail(MPD_ERR_INVALID, "no fused kernel for this configuration");
    if (h->cfg.precision == MPD_FP32_BULK && path == MPD_PATH_PHASES && !is_pp(h->cfg))
        return fail(MPD_ERR_INVALID, "the phase path is FP64 only (and FP32 for the pseudopotential model)");
    if (h->cfg.precision == MPD_FP32_BULK && path == MPD_PATH_PHASES && h->cfg.flux == MPD_FLUX_UW)
        return fail(MPD_ERR_INVALID, "the FP32 pseudopotential step with UW faces has one launch per step");
    if (h->cfg.precision == MPD_FP32_BULK && path == MPD_PATH_PHASES && h->cfg.model == MPD_MODEL_PP_D2V16)
        return fail(MPD_ERR_INVALID, "the FP32 D2V16 step has one launch per step");
    h->path = path;
    return 0;
}

int mpd_set_face_order(mpd_handle *h, int32_t order)
{
    if (!h) return fail(MPD_ERR_INVALID, "null handle");
    if (h->batch) return fail(MPD_ERR_STATE, "handle is a member of a batch");
    if (order != MPD_FACE_ORDER_JACOBI && order != MPD_FACE_ORDER_REFERENCE)
        return fail(MPD_ERR_INVALID, "unknown face order %d", order);
    if (order == MPD_FACE_ORDER_REFERENCE &&
        !(h->cfg.model == MPD_MODEL_PP_D2Q9 && h->cfg.precision == MPD_FP64 && h->cfg.bc_y == MPD_BC_PERIODIC))
        return fail(MPD_ERR_INVALID, "the reference face order is a mode of the FP64 pseudopotential D2Q9 step "
                                     "(periodic mesh)");
    h->C.face_ref = order == MPD_FACE_ORDER_REFERENCE;
    return 0;
}

int mpd_upload_face_psi(mpd_handle *h, const double *psi)
{
    if (!h || !psi) return fail(MPD_ERR_INVALID, "null argument");
    if (h->cfg.model != MPD_MODEL_PP_D2Q9 || h->cfg.precision != MPD_FP64)
        return fail(MPD_ERR_INVALID, "face Psi is state of the FP64 pseudopotential D2Q9 step");
    HIPCHK(hipMemcpyAsync(h->S.fpsi_old, psi, 2 * (size_t)h->S.N * sizeof(double), hipMemcpyHostToDevice,
                          h->stream));
    // the second copies of the periodic faces take the first copies' Psi
    pp_copies_from_first(h->S, h->C, false, true, h->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int mpd_download_face_psi(mpd_handle *h, double *psi)
{
    if (!h || !psi) return fail(MPD_ERR_INVALID, "null argument");
    if (h->cfg.model != MPD_MODEL_PP_D2Q9 || h->cfg.precision != MPD_FP64)
        return fail(MPD_ERR_INVALID, "face Psi is state of the FP64 pseudopotential D2Q9 step");
    HIPCHK(hipMemcpyAsync(psi, h->S.fpsi_old, 2 * (size_t)h->S.N * sizeof(double), hipMemcpyDeviceToHost,
                          h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int mpd_set_tile(mpd_handle *h, int32_t tile)
{
    if (!h) return fail(MPD_ERR_INVALID, "null handle");
    if (h->batch) return fail(MPD_ERR_STATE, "handle is a member of a batch: set the batch's tile");
    if (tile < MPD_TILE_AUTO || tile > MPD_TILE_32X16) return fail(MPD_ERR_INVALID, "unknown tile shape %d", tile);
    if (is_pp(h->cfg) && tile == MPD_TILE_16X16)
        return fail(MPD_ERR_INVALID, "the pseudopotential step has 16x15 and 32x16 tiles");
    if (!is_pp(h->cfg) && h->cfg.precision == MPD_FP64 && tile == MPD_TILE_32X16)
        return fail(MPD_ERR_INVALID, "the FP64 step has 16x15 and 16x16 tiles");
    if (h->cfg.flux == MPD_FLUX_UW && tile != MPD_TILE_AUTO && tile != MPD_TILE_16X15)
        return fail(MPD_ERR_INVALID, "the UW step has 16x15 tiles");
    if (h->cfg.model == MPD_MODEL_PP_D2V16 && tile != MPD_TILE_AUTO && tile != MPD_TILE_16X15)
        return fail(MPD_ERR_INVALID, "the D2V16 step has 16x15 tiles");
    h->tile = tile;
    return 0;
}

int mpd_get_tile(mpd_handle *h, int32_t *tile)
{
    if (!h || !tile) return fail(MPD_ERR_INVALID, "null argument");
    if (h->tile) *tile = h->tile;
    else if (h->cfg.flux == MPD_FLUX_UW) *tile = MPD_TILE_16X15;
    else if (is_pp(h->cfg)) *tile = pp32_tile_auto(h->cfg.nx, h->cfg.ny, num_q(h->cfg), h->cus);
    else if (h->cfg.precision == MPD_FP64) *tile = fused64_tile_auto(h->cfg.nx, h->cfg.ny, h->cus);
    else *tile = fused32_tile_auto(h->cfg.nx, h->cfg.ny, h->cus, fused32_state_bytes(h->cfg));
    return 0;
}

static int alloc_phase_scratch(mpd_handle *h)
{
    DevState &S = h->S;
    if (S.fbp) return 0;
    const size_t Q = (size_t)num_q(h->cfg), N = S.N, QN = Q * N, yfn = (size_t)h->C.nyf * h->cfg.nx;
    bool pp = is_pp(h->cfg), ac = !pp, uw = h->cfg.flux == MPD_FLUX_UW;
    // UW: fluxes of the second copies of the periodic faces (Allen-Cahn; pseudopotential
    // with its copies' state, allocated at create), and the pseudopotential copies' scratch
    const bool pp_copies = S.cbpsi_old != nullptr;
    const size_t nb = (size_t)h->cfg.nx + h->cfg.ny;
    int r;
    if ((r = dalloc(h, &S.hbp, ac ? QN : 0)) || (r = dalloc(h, &S.fbp, QN)) ||
        (r = dalloc(h, &S.hbpx, ac && uw ? QN : 0)) || (r = dalloc(h, &S.hbpy, ac && uw ? QN : 0)) ||
        (r = dalloc(h, &S.fbpx, uw ? QN : 0)) || (r = dalloc(h, &S.fbpy, uw ? QN : 0)) ||
        (r = dalloc(h, &S.hfl, ac ? QN + Q * yfn : 0)) ||
        (r = dalloc(h, &S.ffl, QN + Q * yfn)) ||
        (r = dalloc(h, &S.gm, h->C.walls ? (size_t)MPD_M_COUNT * 2 * h->cfg.nx : 0)) ||
        (r = dalloc(h, &S.ghbp, h->C.walls ? (size_t)Q9 * 2 * h->cfg.nx : 0)) ||
        (r = dalloc(h, &S.gfbp, h->C.walls ? (size_t)Q9 * 2 * h->cfg.nx : 0)) ||
        (r = dalloc(h, &S.cbhl, ac && uw ? Q9 * nb : 0)) ||
        (r = dalloc(h, &S.cbfl, (ac && uw) || pp_copies ? Q9 * nb : 0)) ||
        (r = dalloc(h, &S.cbpsi, pp_copies ? nb : 0)) || (r = dalloc(h, &S.cbscr, pp_copies ? 4 * nb : 0)) ||
        (r = dalloc(h, &S.fpsi, pp ? 2 * N : 0)) || (r = dalloc(h, &S.fscr, pp ? 8 * N : 0)))
        return r;
    return 0;
}

static int take_event(mpd_handle *h, hipEvent_t *e)
{
    if (!h->ev_free.empty()) {
        *e = h->ev_free.back();
        h->ev_free.pop_back();
        return 0;
    }
    HIPCHK(hipEventCreate(e));
    return 0;
}

// Add the elapsed time of every pending event pair to the timing counters
// (waits for the last pair; the stream is usually idle by then).
static int resolve_timing(mpd_handle *h)
{
    for (auto &pr : h->pending) {
        HIPCHK(hipEventSynchronize(pr.second));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, pr.first, pr.second));
        h->acc_ms += ms;
        h->ev_free.push_back(pr.first);
        h->ev_free.push_back(pr.second);
    }
    h->pending.clear();
    return 0;
}

int mpd_upload_state(mpd_handle *h, const double *tilde_h, const double *tilde_f, const double *macro,
                     const double *face_fx, const double *face_fy)
{
    if (!h) return fail(MPD_ERR_INVALID, "null handle");
    if (h->batch) return fail(MPD_ERR_STATE, "handle is a member of a batch: remove it first");
    if (!tilde_f || !macro) return fail(MPD_ERR_INVALID, "tilde_f and macro are required");
    bool ac = !is_pp(h->cfg);
    if (ac && !tilde_h) return fail(MPD_ERR_INVALID, "tilde_h is required for the Allen-Cahn models");
    const size_t N = h->S.N, QN = (size_t)num_q(h->cfg) * N;
    if (ac) HIPCHK(hipMemcpyAsync(h->S.h, tilde_h, QN * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->S.f, tilde_f, QN * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->S.m, macro, (size_t)MPD_M_COUNT * N * sizeof(double), hipMemcpyHostToDevice,
                          h->stream));
    if (!ac) {
        if (face_fx) HIPCHK(hipMemcpyAsync(h->S.ffx, face_fx, 2 * N * sizeof(double), hipMemcpyHostToDevice, h->stream));
        else HIPCHK(hipMemsetAsync(h->S.ffx, 0, 2 * N * sizeof(double), h->stream));
        if (face_fy) HIPCHK(hipMemcpyAsync(h->S.ffy, face_fy, 2 * N * sizeof(double), hipMemcpyHostToDevice, h->stream));
        else HIPCHK(hipMemsetAsync(h->S.ffy, 0, 2 * N * sizeof(double), h->stream));
        // face Psi before the first step: 0, as the reference's initialisers leave it
        // (Ark/Cpp/Preprocess.cpp:337-401 never set it); mpd_upload_face_psi sets it
        HIPCHK(hipMemsetAsync(h->S.fpsi_old, 0, 2 * N * sizeof(double), h->stream));
        // both stored copies of a periodic face start from the uploaded force
        pp_copies_from_first(h->S, h->C, true, true, h->stream);
        HIPCHK(hipGetLastError());
    }
    // one relaxation time for every cell (every pseudopotential initialiser
    // writes pp_tau): the FP32 step then takes it as a constant
    h->tau_u = 0.0;
    if (!ac) {
        const double *tau = macro + (size_t)MPD_M_TAU * N, t0 = tau[0];
        bool same = t0 > 0.0;
        for (size_t n = 1; same && n < N; ++n) same = tau[n] == t0;
        if (same) h->tau_u = t0;
    }
    // Rho_1k starts at 0 in the reference (Ark/Header/Macroscopic.h:44)
    HIPCHK(hipMemsetAsync(h->S.rho1k, 0, N * sizeof(double), h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->uploaded = true;
    h->fused_primed = false;
    h->prime_stage = 0;
    return 0;
}

// Bring the explicit macro arrays up to date (fused path keeps a derived state).
static int sync_macros(mpd_handle *h)
{
    if (h->effective_path == MPD_PATH_FUSED && h->fused_primed) {
        int r = fused_materialize(h->F, h->S, h->C, h->cfg, h->stream);
        if (r) return fail(MPD_ERR_DEVICE, "materialize: %s", hipGetErrorString(hipGetLastError()));
    }
    return 0;
}

static int device_state_ready(mpd_handle *h)
{
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    h->uploaded = true;
    h->fused_primed = false;
    h->prime_stage = 0;
    return 0;
}

int mpd_device_init_drops(mpd_handle *h, int32_t ndrops, const double *cx, const double *cy, const double *radius,
                          int32_t sign, double p0, double mu0)
{
    if (!h) return fail(MPD_ERR_INVALID, "null handle");
    if (h->batch) return fail(MPD_ERR_STATE, "handle is a member of a batch");
    if (is_pp(h->cfg)) return fail(MPD_ERR_INVALID, "drops start the Allen-Cahn models");
    if (ndrops < 1 || ndrops > MAX_DROPS || !cx || !cy || !radius)
        return fail(MPD_ERR_INVALID, "need 1..%d drops", MAX_DROPS);
    DropSet D;
    D.n = ndrops;
    for (int d = 0; d < ndrops; ++d) {
        D.cx[d] = cx[d];
        D.cy[d] = cy[d];
        D.r[d] = radius[d];
    }
    D.width = h->cfg.width;
    devinit_drops(h->S, h->C, D, sign, p0, mu0, h->stream);
    return device_state_ready(h);
}

int mpd_device_init_interface(mpd_handle *h, const double *y0, int32_t sign, double p0)
{
    if (!h) return fail(MPD_ERR_INVALID, "null handle");
    if (h->batch) return fail(MPD_ERR_STATE, "handle is a member of a batch");
    if (is_pp(h->cfg) || !y0) return fail(MPD_ERR_INVALID, "interface needs an Allen-Cahn model and y0[nx]");
    double *d_y0 = nullptr;
    HIPCHK(hipMalloc(&d_y0, sizeof(double) * (size_t)h->cfg.nx));
    hipError_t e = hipMemcpyAsync(d_y0, y0, sizeof(double) * (size_t)h->cfg.nx, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) devinit_interface(h->S, h->C, d_y0, sign, p0, h->cfg.width, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(d_y0);
    if (e != hipSuccess) return fail(MPD_ERR_DEVICE, "interface start: %s", hipGetErrorString(e));
    return device_state_ready(h);
}

}  // extern "C"

// ---- the allocation-free variant of mpd_device_init_interface ----
// The per-column heights are staged in this file-scope device array instead of
// a hipMalloc'ed buffer.  Each GPU holds its own copy at its own address, which
// host code takes from hipGetSymbolAddress (the array's name in host code is a
// host-side stand-in, not a device address).  Every call on one GPU shares that
// GPU's copy: calls that overlap on one GPU race.  Each call overwrites the nx
// values it uses, so nothing needs clearing.
namespace {
constexpr int NOALLOC_Y0_MAX = 16384;  // columns (128 KiB)
__device__ __align__(256) double scratch[NOALLOC_Y0_MAX];
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[MAX_DEVICES] = {};

// address of `scratch` on the current device `dev` (looked up once per device)
int scratch_on(int dev, double **out)
{
    if (dev < 0 || dev >= MAX_DEVICES) return fail(MPD_ERR_DEVICE, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(MPD_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = (double *)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}
}  // namespace

extern "C" {

int mpd_device_init_interface_noalloc(mpd_handle *h, const double *y0, int32_t sign, double p0)
{
    if (!h) return fail(MPD_ERR_INVALID, "null handle");
    if (h->batch) return fail(MPD_ERR_STATE, "handle is a member of a batch");
    if (is_pp(h->cfg) || !y0) return fail(MPD_ERR_INVALID, "interface needs an Allen-Cahn model and y0[nx]");
    if (h->cfg.nx > NOALLOC_Y0_MAX)
        return fail(MPD_ERR_INVALID, "mpd_device_init_interface_noalloc takes at most %d columns (nx = %d); "
                                     "use mpd_device_init_interface", NOALLOC_Y0_MAX, h->cfg.nx);
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    if (cur != h->device) HIPCHK(hipSetDevice(h->device));
    double *d_y0 = nullptr;
    int r = scratch_on(h->device, &d_y0);
    hipError_t e = hipSuccess;
    if (!r) e = hipMemcpyAsync(d_y0, y0, sizeof(double) * (size_t)h->cfg.nx, hipMemcpyHostToDevice, h->stream);
    if (!r && e == hipSuccess) devinit_interface(h->S, h->C, d_y0, sign, p0, h->cfg.width, h->stream);
    if (!r && e == hipSuccess) r = device_state_ready(h);  // waits: the scratch is free again on return
    else if (!r) r = fail(MPD_ERR_DEVICE, "interface start: %s", hipGetErrorString(e));
    if (cur != h->device) (void)hipSetDevice(cur);
    return r;
}

int mpd_device_init_pp_noise(mpd_handle *h, double rho0, double amp, uint64_t seed, double mu0)
{
    if (!h) return fail(MPD_ERR_INVALID, "null handle");
    if (h->batch) return fail(MPD_ERR_STATE, "handle is a member of a batch");
    if (!is_pp(h->cfg)) return fail(MPD_ERR_INVALID, "noise start is for the pseudopotential models");
    devinit_noise(h->S, h->C, rho0, amp, seed, mu0, h->stream);
    h->tau_u = h->C.pp_tau;  // every cell: pp_tau
    return device_state_ready(h);
}

int mpd_device_init_pp_drop(mpd_handle *h, double cx, double cy, double radius, double rho_l, double rho_v,
                            double width, double mu0)
{
    if (!h) return fail(MPD_ERR_INVALID, "null handle");
    if (h->batch) return fail(MPD_ERR_STATE, "handle is a member of a batch");
    if (!is_pp(h->cfg)) return fail(MPD_ERR_INVALID, "SCMP_Drop starts the pseudopotential models");
    if (!(width > 0.0)) return fail(MPD_ERR_INVALID, "width must be positive");
    devinit_pp_drop(h->S, h->C, cx, cy, radius, rho_l, rho_v, width, mu0, h->stream);
    h->tau_u = h->C.pp_tau;
    return device_state_ready(h);
}

int mpd_device_init_pp_flat(mpd_handle *h, double y_lo, double y_hi, double slope, double rho_l, double rho_v,
                            double mu0)
{
    if (!h) return fail(MPD_ERR_INVALID, "null handle");
    if (h->batch) return fail(MPD_ERR_STATE, "handle is a member of a batch");
    if (!is_pp(h->cfg)) return fail(MPD_ERR_INVALID, "SCMP_FlatInterface starts the pseudopotential models");
    devinit_pp_flat(h->S, h->C, y_lo, y_hi, slope, rho_l, rho_v, mu0, h->stream);
    h->tau_u = h->C.pp_tau;
    return device_state_ready(h);
}

int mpd_download_state(mpd_handle *h, double *tilde_h, double *tilde_f, double *macro, double *face_fx,
                       double *face_fy)
{
    if (!h) return fail(MPD_ERR_INVALID, "null handle");
    if (h->batch) HIPCHK(hipStreamSynchronize(h->batch->stream));
    if (!h->uploaded) return fail(MPD_ERR_STATE, "no state uploaded");
    if (int r = sync_macros(h)) return r;
    const size_t N = h->S.N, QN = (size_t)num_q(h->cfg) * N;
    bool ac = !is_pp(h->cfg);
    if (tilde_h && ac) HIPCHK(hipMemcpyAsync(tilde_h, h->S.h, QN * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if (tilde_f) HIPCHK(hipMemcpyAsync(tilde_f, h->S.f, QN * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if (macro)
        HIPCHK(hipMemcpyAsync(macro, h->S.m, (size_t)MPD_M_COUNT * N * sizeof(double), hipMemcpyDeviceToHost,
                              h->stream));
    if (!ac && face_fx) HIPCHK(hipMemcpyAsync(face_fx, h->S.ffx, 2 * N * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if (!ac && face_fy) HIPCHK(hipMemcpyAsync(face_fy, h->S.ffy, 2 * N * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int mpd_step(mpd_handle *h, int64_t nsteps)
{
    if (!h) return fail(MPD_ERR_INVALID, "null handle");
    if (h->batch) return fail(MPD_ERR_STATE, "handle is a member of a batch: step the batch");
    if (!h->uploaded) return fail(MPD_ERR_STATE, "no state uploaded");
    if (nsteps < 0) return fail(MPD_ERR_INVALID, "nsteps < 0");
    if (nsteps == 0) return 0;
    int path = h->path;
    if (path == MPD_PATH_AUTO) path = fused_supported(h->cfg) ? MPD_PATH_FUSED : MPD_PATH_PHASES;
    // FP32-bulk pseudopotential: the phase path is the same FP32 state stepped
    // with one launch per phase instead of one per step
    const bool pp32_phases = h->cfg.precision == MPD_FP32_BULK && is_pp(h->cfg) && path == MPD_PATH_PHASES;
    if (pp32_phases) path = MPD_PATH_FUSED;
    if (h->cfg.precision == MPD_FP32_BULK && path != MPD_PATH_FUSED)
        return fail(MPD_ERR_INVALID, "FP32-bulk needs the fused path");
    if (path != h->effective_path) {
        // leaving the fused path: write its derived state back as explicit macros
        if (h->fused_primed)
            if (int r = sync_macros(h)) return r;
        h->fused_primed = false;
        h->prime_stage = 0;
    }
    h->effective_path = path;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    // an error return hands the timing pair back to the pool
    auto bail = [&](int r) {
        if (e0) h->ev_free.push_back(e0);
        if (e1) h->ev_free.push_back(e1);
        return r;
    };
    // Without per-launch timing one event pair brackets the timed launches.
    // On the fused path the start event follows the first step kernel when
    // there are two or more: an event ahead of the first kernel is stamped as
    // soon as the idle stream reaches it, before the host has enqueued that
    // kernel, so the pair would also time the host's enqueue latency.
    // `timed` counts the launches inside the pair.
    int64_t timed = 0;
    if (!h->per_launch) {
        if (int r = take_event(h, &e0)) return r;
        if (int r = take_event(h, &e1)) return bail(r);
    }
    auto mark = [&]() -> int {
        if (e0)
            if (hipError_t e = hipEventRecord(e0, h->stream); e != hipSuccess)
                return bail(fail(MPD_ERR_DEVICE, "hipEventRecord: %s", hipGetErrorString(e)));
        return 0;
    };
    if (path == MPD_PATH_PHASES) {
        if (int r = alloc_phase_scratch(h)) return bail(r);
        if (int r = mark()) return r;
        timed = nsteps;
        for (int64_t s = 0; s < nsteps; ++s) {
            phase_step(h->S, h->C, h->stream);
            // this step's face Psi becomes the previous step's for the next one
            if (is_pp(h->cfg)) std::swap(h->S.fpsi, h->S.fpsi_old);
            if (h->S.cbpsi) std::swap(h->S.cbpsi, h->S.cbpsi_old);
        }
    } else {
        if (!h->F.ready) {
            int r = fused_alloc(h->F, h->S, h->C, h->cfg);
            if (r) return bail(fail(MPD_ERR_NOMEM, "fused path allocation failed"));
        }
        h->F.pp32.phases = pp32_phases;
        h->F.cus = h->cus;
        if (is_pp(h->cfg)) {
            h->F.pp32.tile = h->tile;
            h->F.pp32.tau_u = h->tau_u;
        }
        else if (h->cfg.precision == MPD_FP32_BULK)
            h->F.tile = h->tile ? h->tile
                                : fused32_tile_auto(h->cfg.nx, h->cfg.ny, h->cus, fused32_state_bytes(h->cfg));
        else
            h->F.tile = h->tile ? h->tile : fused64_tile_auto(h->cfg.nx, h->cfg.ny, h->cus);
        int64_t s = 0;
        if (!h->fused_primed && !fused_needs_priming(h->cfg)) {
            // explicit state carried as is (FP32-bulk pseudopotential)
            fused_convert(h->F, h->S, h->C, 0, h->stream);
            h->fused_primed = true;
        }
        if (!h->fused_primed) {
            // The derived state needs two explicit steps: the first step after an
            // upload uses the uploaded sources (not derivable from phi), and the
            // momentum update of a step uses the sources of the step before.
            if (int r = alloc_phase_scratch(h)) return bail(r);
            for (; s < nsteps && h->prime_stage < 2; ++s) {
                if (h->prime_stage == 0) fused_convert(h->F, h->S, h->C, 0, h->stream);
                phase_step(h->S, h->C, h->stream);
                fused_convert(h->F, h->S, h->C, h->prime_stage + 1, h->stream);
                ++h->prime_stage;
            }
            if (h->prime_stage == 2) h->fused_primed = true;
        }
        if (h->per_launch) {
            // bracket each step kernel with events, in chunks of the pool size
            const int64_t pool = 512;
            if (h->evs.empty()) {
                h->evs.resize(2 * pool);
                for (auto &e : h->evs) HIPCHK(hipEventCreate(&e));
            }
            double ms_sum = 0.0;
            while (s < nsteps) {
                int64_t k = 0;
                for (; k < pool && s < nsteps; ++k, ++s) {
                    HIPCHK(hipEventRecord(h->evs[2 * k], h->stream));
                    fused_step(h->F, h->S, h->C, h->cfg, 0, h->stream);
                    HIPCHK(hipEventRecord(h->evs[2 * k + 1], h->stream));
                }
                HIPCHK(hipEventSynchronize(h->evs[2 * k - 1]));
                for (int64_t a = 0; a < k; ++a) {
                    float ms = 0.f;
                    HIPCHK(hipEventElapsedTime(&ms, h->evs[2 * a], h->evs[2 * a + 1]));
                    ms_sum += ms;
                }
                timed += k;
            }
            h->acc_ms += ms_sum;
            h->launches += timed;
            HIPCHK(hipGetLastError());
            return 0;
        }
        if (nsteps - s >= 2) {
            fused_step(h->F, h->S, h->C, h->cfg, 0, h->stream);
            ++s;
        }
        if (int r = mark()) return r;
        timed = nsteps - s;
        for (; s < nsteps; ++s) fused_step(h->F, h->S, h->C, h->cfg, 0, h->stream);
    }
    HIPCHK(hipGetLastError());
    if (e1) {
        if (timed == 0) {
            // every step was a priming step: nothing to time
            h->ev_free.push_back(e0);
            h->ev_free.push_back(e1);
            return 0;
        }
        HIPCHK(hipEventRecord(e1, h->stream));
        h->pending.emplace_back(e0, e1);
        // bound the pool: a caller that never synchronises waits here now and then
        if (h->pending.size() >= 256)
            if (int r = resolve_timing(h)) return r;
    }
    h->launches += timed;
    return 0;
}

int mpd_kernel_timing(mpd_handle *h, double *avg_ms, int64_t *launches)
{
    if (!h) return fail(MPD_ERR_INVALID, "null handle");
    if (int r = resolve_timing(h)) return r;
    if (avg_ms) *avg_ms = h->launches ? h->acc_ms / (double)h->launches : 0.0;
    if (launches) *launches = h->launches;
    h->acc_ms = 0.0;
    h->launches = 0;
    return 0;
}

int mpd_set_kernel_timing(mpd_handle *h, int enable)
{
    if (!h) return fail(MPD_ERR_INVALID, "null handle");
    if (int r = resolve_timing(h)) return r;
    h->per_launch = enable != 0;
    h->acc_ms = 0.0;
    h->launches = 0;
    return 0;
}

int mpd_synchronize(mpd_handle *h)
{
    if (!h) return fail(MPD_ERR_INVALID, "null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    return resolve_timing(h);
}

int mpd_device_bytes(const mpd_handle *h, int64_t *bytes)
{
    if (!h || !bytes) return fail(MPD_ERR_INVALID, "null argument");
    *bytes = h->bytes + fused_bytes(h->F);
    return 0;
}

int mpd_reductions(mpd_handle *h, int update_residual, mpd_sums *out)
{
    if (!h || !out) return fail(MPD_ERR_INVALID, "null argument");
    if (h->batch) HIPCHK(hipStreamSynchronize(h->batch->stream));
    if (!h->uploaded) return fail(MPD_ERR_STATE, "no state uploaded");
    if (int r = sync_macros(h)) return r;
    hipLaunchKernelGGL(k_reduce, dim3(h->red_blocks), dim3(RED_THREADS), 0, h->stream, h->S.m, h->S.rho1k,
                       h->S.N, update_residual, h->d_part);
    HIPCHK(hipGetLastError());
    std::vector<double> part((size_t)h->red_blocks * RED_FIELDS);
    HIPCHK(hipMemcpyAsync(part.data(), h->d_part, part.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    double s[7] = {0}, c[7] = {0}, vmax = 0.0, bad = 0.0;
    for (int b = 0; b < h->red_blocks; ++b) {
        for (int k = 0; k < 7; ++k) {
            double x = part[(size_t)b * RED_FIELDS + k], t = s[k] + x;
            if (std::fabs(s[k]) >= std::fabs(x)) c[k] += (s[k] - t) + x;
            else c[k] += (x - t) + s[k];
            s[k] = t;
        }
        vmax = std::fmax(vmax, part[(size_t)b * RED_FIELDS + 7]);
        bad += part[(size_t)b * RED_FIELDS + 8];
    }
    out->sum_rho = s[0] + c[0];
    out->sum_rho_u = s[1] + c[1];
    out->sum_rho_v = s[2] + c[2];
    out->sum_phi = s[3] + c[3];
    out->kinetic_energy = s[4] + c[4];
    out->max_speed = vmax;
    out->residual = std::sqrt((s[5] + c[5]) / ((s[6] + c[6]) + 1.0E-30));
    out->nonfinite = (int64_t)bad;
    if (bad > 0) return fail(MPD_ERR_NONFINITE, "%lld cells with NaN/Inf or rho <= 0", (long long)bad);
    return 0;
}

// ---------------- batched ensembles ----------------

static void batch_drop_graph(mpd_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

int mpd_batch_create(mpd_batch **out)
{
    if (!out) return fail(MPD_ERR_INVALID, "null output pointer");
    *out = nullptr;
    mpd_batch *b = new (std::nothrow) mpd_batch();
    if (!b) return fail(MPD_ERR_NOMEM, "host allocation failed");
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    b->plan = batch32_new();
    if (e != hipSuccess || !b->plan) {
        mpd_batch_destroy(b);
        return fail(MPD_ERR_DEVICE, "batch set-up failed: %s", hipGetErrorString(e));
    }
    *out = b;
    return 0;
}

int mpd_batch_destroy(mpd_batch *b)
{
    if (!b) return 0;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (mpd_handle *h : b->members) h->batch = nullptr;
    batch_drop_graph(b);
    batch32_free(b->plan);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return 0;
}

int mpd_batch_add(mpd_batch *b, mpd_handle *h)
{
    if (!b || !h) return fail(MPD_ERR_INVALID, "null argument");
    if (h->batch) return fail(MPD_ERR_STATE, "handle already belongs to a batch");
    if (h->cfg.precision != MPD_FP32_BULK || !fused_supported(h->cfg) || is_pp(h->cfg) ||
        h->cfg.flux != MPD_FLUX_CD)
        return fail(MPD_ERR_INVALID, "batches take FP32-bulk cases of the fused path (Allen-Cahn, CD)");
    if (!h->fused_primed || h->effective_path != MPD_PATH_FUSED)
        return fail(MPD_ERR_STATE, "step the case at least twice on the fused path before adding it");
    if (!b->members.empty()) {
        const mpd_config &c0 = b->members[0]->cfg;
        if (c0.model != h->cfg.model || c0.bc_y != h->cfg.bc_y)
            return fail(MPD_ERR_INVALID, "all cases of a batch share the model and the y boundary");
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    b->members.push_back(h);
    h->batch = b;
    b->dirty = true;
    b->tune_pending = true;
    return 0;
}

int mpd_batch_remove(mpd_batch *b, mpd_handle *h)
{
    if (!b || !h) return fail(MPD_ERR_INVALID, "null argument");
    if (h->batch != b) return fail(MPD_ERR_STATE, "handle is not a member of this batch");
    HIPCHK(hipStreamSynchronize(b->stream));
    for (size_t k = 0; k < b->members.size(); ++k)
        if (b->members[k] == h) {
            b->members.erase(b->members.begin() + (long)k);
            break;
        }
    h->batch = nullptr;
    b->dirty = true;
    b->tune_pending = true;
    return 0;
}

int mpd_batch_size(const mpd_batch *b, int32_t *ncases, int64_t *ncells)
{
    if (!b) return fail(MPD_ERR_INVALID, "null batch");
    int64_t cells = 0;
    for (mpd_handle *h : b->members) cells += h->S.N;
    if (ncases) *ncases = (int32_t)b->members.size();
    if (ncells) *ncells = cells;
    return 0;
}

static bool batch_nee(const mpd_batch *b)
{
    return !b->members.empty() && b->members[0]->cfg.bc_y == MPD_BC_WALL_NEE;
}

// device tables of the current members with tiles of shape b->tile (no graph)
static int batch_tables(mpd_batch *b)
{
    batch_drop_graph(b);
    std::vector<BatchMember> m(b->members.size());
    for (size_t k = 0; k < b->members.size(); ++k) {
        mpd_handle *h = b->members[k];
        m[k].C = &h->C;
        m[k].N = h->S.N;
        m[k].pl[0] = fused32_planes(h->F, 0);
        m[k].pl[1] = fused32_planes(h->F, 1);
        m[k].cur = h->F.cur;
    }
    if (batch32_plan(b->plan, m.data(), (int)m.size(), b->tile, b->stream))
        return fail(MPD_ERR_DEVICE, "batch tables: %s", hipGetErrorString(hipGetLastError()));
    return 0;
}

static int batch_rebuild(mpd_batch *b)
{
    if (int r = batch_tables(b)) return r;
    // one graph = two steps (in->out, out->in), so the planes are back in place after each replay
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    batch32_step(b->plan, 0, b->stream);
    batch32_step(b->plan, 1, b->stream);
    HIPCHK(hipStreamEndCapture(b->stream, &b->graph));
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->dirty = false;
    return 0;
}

// Tile-shape autotuner: TUNE_STEPS steps of the batch with each tile shape;
// of each, the first TUNE_WARM warm the new tables and code, the rest are
// timed with events on the batch stream.  The fastest shape is kept until the
// members change.  The shapes give the same state bit for bit with periodic
// and bounce-back walls (each cell and face is computed by the same code), so
// the tuning steps are ordinary steps of the run.  Even counts leave every
// member's current plane set in place.
constexpr int TUNE_STEPS = 8, TUNE_WARM = 2, NSHAPES = 3;

static int batch_tune(mpd_batch *b)
{
    for (int c = 0; c < NSHAPES; ++c) {
        b->tile = MPD_TILE_16X15 + c;
        if (int r = batch_tables(b)) return r;
        HIPCHK(hipEventRecord(b->ev1, b->stream));
        for (int s = 0; s < TUNE_STEPS; ++s) {
            if (s == TUNE_WARM) HIPCHK(hipEventRecord(b->ev0, b->stream));
            batch32_step(b->plan, s & 1, b->stream);
        }
        HIPCHK(hipGetLastError());
        hipEvent_t e_all = nullptr;
        HIPCHK(hipEventCreate(&e_all));
        hipError_t e = hipEventRecord(e_all, b->stream);
        if (e == hipSuccess) e = hipEventSynchronize(e_all);
        float ms = 0.f, ms_all = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, b->ev0, e_all);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms_all, b->ev1, e_all);
        (void)hipEventDestroy(e_all);
        if (e != hipSuccess) return fail(MPD_ERR_DEVICE, "batch tuning: %s", hipGetErrorString(e));
        b->tune_ms[c] = ms / (TUNE_STEPS - TUNE_WARM);
        b->acc_ms += ms_all;
        b->steps += TUNE_STEPS;
    }
    int best = 0;
    for (int c = 1; c < NSHAPES; ++c)
        if (b->tune_ms[c] < b->tune_ms[best]) best = c;
    b->tile = MPD_TILE_16X15 + best;
    b->tune_pending = false;
    b->dirty = true;
    return 0;
}

int mpd_batch_set_tile(mpd_batch *b, int32_t tile)
{
    if (!b) return fail(MPD_ERR_INVALID, "null batch");
    if (tile < MPD_TILE_AUTO || tile > MPD_TILE_32X16) return fail(MPD_ERR_INVALID, "unknown tile shape %d", tile);
    HIPCHK(hipStreamSynchronize(b->stream));
    b->tile_req = tile;
    b->tune_pending = tile == MPD_TILE_AUTO;
    b->dirty = true;
    return 0;
}

int mpd_batch_tile(const mpd_batch *b, int32_t *tile, double ms_per_step[3])
{
    if (!b) return fail(MPD_ERR_INVALID, "null batch");
    if (tile) *tile = b->tile_req ? b->tile_req : b->tile;
    if (ms_per_step)
        for (int c = 0; c < NSHAPES; ++c) ms_per_step[c] = b->tune_ms[c];
    return 0;
}

int mpd_batch_step(mpd_batch *b, int64_t nsteps)
{
    if (!b) return fail(MPD_ERR_INVALID, "null batch");
    if (nsteps < 0) return fail(MPD_ERR_INVALID, "nsteps < 0");
    if (nsteps == 0 || b->members.empty()) return 0;
    if (b->tile_req) {
        if (b->tile != b->tile_req) b->dirty = true;
        b->tile = b->tile_req;
    } else if (batch_nee(b)) {
        if (b->tile != MPD_TILE_16X15) b->dirty = true;
        b->tile = MPD_TILE_16X15;   // NEE: the shapes differ by FP32 rounding; keep results independent of timing
    } else if (b->tune_pending && nsteps >= NSHAPES * TUNE_STEPS) {
        if (int r = batch_tune(b)) return r;
        nsteps -= NSHAPES * TUNE_STEPS;
        if (nsteps == 0) return 0;
    }
    if (b->dirty)
        if (int r = batch_rebuild(b)) return r;
    HIPCHK(hipEventRecord(b->ev0, b->stream));
    for (int64_t s = 0; s + 1 < nsteps; s += 2) HIPCHK(hipGraphLaunch(b->exec, b->stream));
    if (nsteps % 2) {
        batch32_step(b->plan, 0, b->stream);
        HIPCHK(hipGetLastError());
        for (mpd_handle *h : b->members) h->F.cur = 1 - h->F.cur;
        b->dirty = true;
    }
    HIPCHK(hipEventRecord(b->ev1, b->stream));
    HIPCHK(hipEventSynchronize(b->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->acc_ms += ms;
    b->steps += nsteps;
    return 0;
}

int mpd_batch_timing(mpd_batch *b, double *avg_ms, int64_t *steps)
{
    if (!b) return fail(MPD_ERR_INVALID, "null batch");
    if (avg_ms) *avg_ms = b->steps ? b->acc_ms / (double)b->steps : 0.0;
    if (steps) *steps = b->steps;
    b->acc_ms = 0.0;
    b->steps = 0;
    return 0;
}

}  // extern "C"
