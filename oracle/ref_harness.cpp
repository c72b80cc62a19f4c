// State-dump hook linked into the reference solver built from its own sources.
//
// Test infrastructure only (a reference build to compare against, never shipped): the recipe in
// oracle/Makefile.ref compiles the reference's unmodified translation units
// and links this file with `--wrap` on the mangled symbols of `Update_phi_Eq` and `Output_Flowfield`.
//
// * `Output_Flowfield` is the first call made by `DUGKS2DSolver()`
//   (Ark/Cpp/DUGKS2DSolver.cpp:150); the wrapper marks the start of the time
//   loop and forwards to the real writer only when MPD_REF_FLOWFIELD=1.
// * `Update_phi_Eq(CellArray[0])` is the first call of every time step
//   (Ark/Cpp/DUGKS2DSolver.cpp:175-183).  When it fires for cell 0 inside the
//   loop, `k` complete steps have been taken, and the whole cell state is
//   written out before the step proceeds.
//
// Environment: MPD_REF_DUMP_STEPS="0,1,2,10" (steps to dump), MPD_REF_TIME_STEPS
// (steps at whose start a steady-clock timestamp goes to stderr), MPD_REF_MAX_STEP
// (exit(0) when reached), MPD_REF_DUMP_DIR (output directory),
// MPD_REF_WRITERS=1 (at each dumped step also call the reference's own
// Output_Flowfield, Output_Diag, Output_MidX and Output_MidY, Ark/Cpp/Output.cpp,
// so their files can serve as format fixtures).
//
// Dump layout (little-endian): int32 nx, ny, nfields, then nfields planes of
// nx*ny float64 in structured order idx = j*nx + i, with (i,j) taken from the
// reference's own CarCellArray[i+1][j+1] (Ark/Cpp/Mesh/MeshCartesian.cpp:369-408).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <string>
#include <vector>
#include "DUGKSDeclaration.h"

// --wrap acts on linker symbols, so the wrappers use the Itanium-mangled names
// of `void Update_phi_Eq(Cell_2D&)` and `void Output_Flowfield(double const&, int)`.
extern "C" void __real__Z13Update_phi_EqR7Cell_2D(Cell_2D &cell);
extern "C" void __real__Z16Output_FlowfieldRKdi(double const &t, int step);
void Output_Diag(int step);
void Output_MidX(int step);
void Output_MidY(int step);

namespace {
bool g_started = false;
int g_steps_seen = -1;
std::set<int> g_dump;
std::set<int> g_time;  // steps at whose start a timestamp is printed
int g_max_step = -1;
std::string g_dir = ".";

void parse_env()
{
    static bool done = false;
    if (done) return;
    done = true;
    if (const char *s = std::getenv("MPD_REF_DUMP_STEPS")) {
        std::string v(s);
        size_t p = 0;
        while (p < v.size()) {
            size_t q = v.find(',', p);
            if (q == std::string::npos) q = v.size();
            if (q > p) g_dump.insert(std::atoi(v.substr(p, q - p).c_str()));
            p = q + 1;
        }
    }
    if (const char *s = std::getenv("MPD_REF_TIME_STEPS")) {
        std::string v(s);
        size_t p = 0;
        while (p < v.size()) {
            size_t q = v.find(',', p);
            if (q == std::string::npos) q = v.size();
            if (q > p) g_time.insert(std::atoi(v.substr(p, q - p).c_str()));
            p = q + 1;
        }
    }
    if (const char *s = std::getenv("MPD_REF_MAX_STEP")) g_max_step = std::atoi(s);
    if (const char *s = std::getenv("MPD_REF_DUMP_DIR")) g_dir = s;
}

void dump_cells(int k)
{
    const int nx = Nx, ny = Ny;
    std::vector<std::vector<double>> planes;
    std::vector<std::string> names;
    auto add = [&](const char *name) {
        names.push_back(name);
        planes.emplace_back(static_cast<size_t>(nx) * ny, 0.0);
        return planes.size() - 1;
    };
#ifdef _ARK_ALLENCAHN_FLIP
    size_t ph = add("h");
    for (int q = 1; q < DV_Qv; ++q) add("h");
#endif
    size_t pf = add("f");
    for (int q = 1; q < DV_Qv; ++q) add("f");
    const char *scal[] = {"Phi", "Rho", "U", "V", "p", "Phi_x", "Phi_y", "Rho_x", "Rho_y",
                          "Fx", "Fy", "laplacianPhi", "prevPhiU", "prevPhiV", "Psi", "scG",
                          "Mu", "ftau", "faBP", "fbBP", "fcBP", "haBP", "hbBP", "hcBP",
                          "xc", "yc"};
    size_t ps = planes.size();
    for (const char *n : scal) add(n);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            Cell_2D const &c = *CarCellArray[i + 1][j + 1];
            size_t idx = static_cast<size_t>(j) * nx + i;
            for (int q = 0; q < DV_Qv; ++q) {
#ifdef _ARK_ALLENCAHN_FLIP
                planes[ph + q][idx] = c.h.Tilde[0][q];
#endif
                planes[pf + q][idx] = c.f.Tilde[0][q];
            }
            MacroQuantity const &m = c.MsQ();
            double v[] = {m.Phi, m.Rho, m.U, m.V, m.p, m.Phi_x, m.Phi_y, m.Rho_x, m.Rho_y,
                          m.Fx, m.Fy, m.laplacianPhi, m.prevPhiU, m.prevPhiV, m.Psi, m.scG,
                          m.Mu, c.f.tau, c.f.aBP, c.f.bBP, c.f.cBP,
#ifdef _ARK_ALLENCAHN_FLIP
                          c.h.aBP, c.h.bBP, c.h.cBP,
#else
                          0.0, 0.0, 0.0,
#endif
                          c.xc, c.yc};
            for (size_t s = 0; s < sizeof(v) / sizeof(v[0]); ++s) planes[ps + s][idx] = v[s];
        }
    char path[4096];
    std::snprintf(path, sizeof(path), "%s/cells_step%d.bin", g_dir.c_str(), k);
    FILE *fp = std::fopen(path, "wb");
    if (!fp) { std::perror(path); std::exit(3); }
    int hdr[3] = {nx, ny, static_cast<int>(planes.size())};
    std::fwrite(hdr, sizeof(int), 3, fp);
    for (auto &p : planes) std::fwrite(p.data(), sizeof(double), p.size(), fp);
    std::fclose(fp);

    // Face state (one record per face in FaceArray order): geometry plus the
    // persistent face macros the pseudopotential model carries between steps.
    std::snprintf(path, sizeof(path), "%s/faces_step%d.bin", g_dir.c_str(), k);
    fp = std::fopen(path, "wb");
    if (!fp) { std::perror(path); std::exit(3); }
    int fh[2] = {Faces, 12};
    std::fwrite(fh, sizeof(int), 2, fp);
    for (int n = 0; n < Faces; ++n) {
        Face_2D const &f = FaceArray[n];
        MacroQuantity const &m = f.MsQh();
        double v[12] = {f.xf, f.yf, f.Vx, f.Vy, static_cast<double>(f.bc_type),
                        m.Rho, m.U, m.V, m.p, m.Psi, m.Fx, m.Fy};
        std::fwrite(v, sizeof(double), 12, fp);
    }
    std::fclose(fp);
    if (const char *s = std::getenv("MPD_REF_WRITERS"))
        if (s[0] == '1') {
            const double t = k * dt;
            __real__Z16Output_FlowfieldRKdi(t, k);
            Output_Diag(k);
            Output_MidX(k);
            Output_MidY(k);
        }
    std::fprintf(stderr, "[ref_harness] dumped step %d\n", k);
}

// Face topology (MPD_REF_TOPO=1): topo.bin = int32 Faces, InteriorFaceNum,
// PeriodicFaceNum, WallFaceNum, then per face object in FaceArray order 9 int32
// faceFaces indices (-1 = none), the shadowF index, bc_type, then 6 float64
// (xf, yf, owner xc, owner yc, neigh xc, neigh yc); then the FaceArray index of
// every InteriorFaceA, PeriodicFaceA and WallFaceA entry, i.e. the order in which
// one thread visits the faces (Ark/Cpp/DUGKS2DSolver.cpp:219-250).
void dump_topology()
{
    char path[4096];
    std::snprintf(path, sizeof(path), "%s/topo.bin", g_dir.c_str());
    FILE *fp = std::fopen(path, "wb");
    if (!fp) { std::perror(path); std::exit(3); }
    auto index = [](Face_2D const *f) { return f ? static_cast<int>(f - FaceArray) : -1; };
    int hdr[4] = {Faces, InteriorFaceNum, PeriodicFaceNum, WallFaceNum};
    std::fwrite(hdr, sizeof(int), 4, fp);
    for (int n = 0; n < Faces; ++n) {
        Face_2D const &f = FaceArray[n];
        int iv[11];
        for (int k = 0; k < 9; ++k) iv[k] = index(f.faceFaces[k]);
        iv[9] = index(f.shadowF);
        iv[10] = f.bc_type;
        double dv[6] = {f.xf, f.yf, f.owner ? f.owner->xc : 0.0, f.owner ? f.owner->yc : 0.0,
                        f.neigh ? f.neigh->xc : 0.0, f.neigh ? f.neigh->yc : 0.0};
        std::fwrite(iv, sizeof(int), 11, fp);
        std::fwrite(dv, sizeof(double), 6, fp);
    }
    for (int n = 0; n < InteriorFaceNum; ++n) { int v = index(InteriorFaceA[n]); std::fwrite(&v, sizeof(int), 1, fp); }
    for (int n = 0; n < PeriodicFaceNum; ++n) { int v = index(PeriodicFaceA[n]); std::fwrite(&v, sizeof(int), 1, fp); }
    for (int n = 0; n < WallFaceNum; ++n) { int v = index(WallFaceA[n]); std::fwrite(&v, sizeof(int), 1, fp); }
    std::fclose(fp);
}
}  // namespace

extern "C" void __wrap__Z16Output_FlowfieldRKdi(double const &t, int step)
{
    parse_env();
    if (!g_started)
        if (const char *s = std::getenv("MPD_REF_TOPO"))
            if (s[0] == '1') dump_topology();
    g_started = true;
    if (const char *s = std::getenv("MPD_REF_FLOWFIELD"))
        if (s[0] == '1') __real__Z16Output_FlowfieldRKdi(t, step);
}

extern "C" void __wrap__Z13Update_phi_EqR7Cell_2D(Cell_2D &cell)
{
    if (g_started && &cell == &CellArray[0]) {
        ++g_steps_seen;
        if (g_time.count(g_steps_seen)) {
            const double t = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
            std::fprintf(stderr, "[ref_harness] step %d t %.6f\n", g_steps_seen, t);
        }
        if (g_dump.count(g_steps_seen)) dump_cells(g_steps_seen);
        if (g_max_step >= 0 && g_steps_seen >= g_max_step) std::exit(0);
    }
    __real__Z13Update_phi_EqR7Cell_2D(cell);
}
