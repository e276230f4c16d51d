// mps_capi.hip -- the C ABI (include/mps_amd.h): scene I/O, the device
// context, the step driver and the per-stage entry points.
//
// Host code only; kernels live in mps_sort.hip and mps_sweeps.hip.  The step
// is enqueued on one stream with no host synchronisation inside it: dt, time,
// n0 and the |u| maximum stay in device memory (DevScalars).
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/mps_amd.h"
#include "mps_kernels.hpp"
#include "mps_ctx.hpp"

using namespace mps;

#define HIPOK(expr)                                     \
    do {                                                \
        hipError_t e_ = (expr);                         \
        if (e_ != hipSuccess) {                         \
            fprintf(stderr, "mps_amd: %s failed: %s\n", #expr, hipGetErrorString(e_)); \
            return MPS_ERR_HIP;                         \
        }                                               \
    } while (0)

// ============================================================ scene I/O ===

extern "C" int mps_params_default(mps_params* p)
{
    if (!p) return MPS_ERR_ARG;
    memset(p, 0, sizeof(*p));
    p->dim = 3;
    p->precision = 0;
    p->rho = 1000.0; p->nu = 1e-6; p->dp = 0.006;
    p->re_n_mult = 2.1; p->re_lap_mult = 2.1; p->beta = 0.97; p->cs = 10.0; p->cfl = 0.1;
    p->col_mult = 0.8; p->col_rate = 2.0; p->dt_max = 1e-4; p->dt_min_mult = 1e-3;
    p->tt = 0.5; p->tout = 0.001; p->max_dis = 0.001;
    p->g[0] = 0.0; p->g[1] = 0.0; p->g[2] = -9.81;
    return MPS_OK;
}

static const char* kLabels[16] = {
    "DIMENTION", "FLUID RHO", "FLUID DYNAMICS VISCOSITY", "INITIAL PARTICLE DISTANCE(DP)",
    "R0(times of DP)", "RLAP(times of DP)", "BETA", "NUMERICAL SPEED OF SOUND", "CFL NUMBER",
    "COLLISION DISTANCE(times of DP)", "COLLISION RATE", "MAX TIME STEP",
    "MIN TIME STEP(times of MAX)", "TT TOTAL CALCULATION TIME", "TOUT TIME STEP OF OUTPUT",
    "MAXIMUM DISPLACEMENT OF THE GEOMETRY"};

static double* param_slot(mps_params* p, int k)
{
    double* s[16] = {nullptr, &p->rho, &p->nu, &p->dp, &p->re_n_mult, &p->re_lap_mult, &p->beta,
                     &p->cs, &p->cfl, &p->col_mult, &p->col_rate, &p->dt_max, &p->dt_min_mult,
                     &p->tt, &p->tout, &p->max_dis};
    return s[k];
}

// input.dat: 16 "label line / value line" pairs in a fixed order (MPS.cpp:62-223).
extern "C" int mps_load_case(const char* path, mps_params* p)
{
    if (!path || !p) return MPS_ERR_ARG;
    FILE* f = fopen(path, "r");
    if (!f) return MPS_ERR_IO;
    mps_params_default(p);
    char buf[256];
    for (int k = 0; k < 16; k++) {
        if (!fgets(buf, sizeof buf, f) || !fgets(buf, sizeof buf, f)) { fclose(f); return MPS_ERR_IO; }
        if (k == 0) {
            if (sscanf(buf, "%d", &p->dim) != 1) { fclose(f); return MPS_ERR_IO; }
        } else if (sscanf(buf, "%lf", param_slot(p, k)) != 1) {
            fclose(f);
            return MPS_ERR_IO;
        }
    }
    fclose(f);
    return (p->dim == 2 || p->dim == 3) ? MPS_OK : MPS_ERR_IO;
}

extern "C" int mps_write_case(const char* path, const mps_params* p)
{
    if (!path || !p) return MPS_ERR_ARG;
    FILE* f = fopen(path, "w");
    if (!f) return MPS_ERR_IO;
    mps_params q = *p;
    for (int k = 0; k < 16; k++) {
        if (k == 0) fprintf(f, "%s\n%d\n", kLabels[k], q.dim);
        else fprintf(f, "%s\n%.17g\n", kLabels[k], *param_slot(&q, k));
    }
    fclose(f);
    return MPS_OK;
}

// GEO.dat: "np nb2 nb1", all positions, all velocities, all pressures (MPS.cpp:236-281).
extern "C" int mps_geo_read(const char* path, int as_float, int* np, int* nb2, int* nb1, double* pos,
                            double* vel, double* press)
{
    if (!path || !np || !nb2 || !nb1) return MPS_ERR_ARG;
    FILE* f = fopen(path, "r");
    if (!f) return MPS_ERR_IO;
    int n, a, b;
    if (fscanf(f, "%d %d %d", &n, &a, &b) != 3 || n < 0 || a < 0 || b < 0 || a + b > n) {
        fclose(f);
        return MPS_ERR_IO;
    }
    *np = n; *nb2 = a; *nb1 = b;
    if (!pos) { fclose(f); return MPS_OK; }
    char tok[128];
    auto next = [&](double* out) -> bool {
        if (fscanf(f, "%127s", tok) != 1) return false;
        char* end = nullptr;
        *out = as_float ? (double)strtof(tok, &end) : strtod(tok, &end);
        return end && *end == 0;
    };
    for (long i = 0; i < 3L * n; i++)
        if (!next(&pos[i])) { fclose(f); return MPS_ERR_IO; }
    for (long i = 0; i < 3L * n; i++) {
        double t;
        if (!next(&t)) { fclose(f); return MPS_ERR_IO; }
        if (vel) vel[i] = t;
    }
    for (long i = 0; i < n; i++) {
        double t;
        if (!next(&t)) { fclose(f); return MPS_ERR_IO; }
        if (press) press[i] = t;
    }
    fclose(f);
    return MPS_OK;
}

extern "C" int mps_geo_write(const char* path, int np, int nb2, int nb1, const double* pos,
                             const double* vel, const double* press)
{
    if (!path || np < 0 || !pos) return MPS_ERR_ARG;
    FILE* f = fopen(path, "w");
    if (!f) return MPS_ERR_IO;
    fprintf(f, "%d %d %d\n", np, nb2, nb1);
    for (long i = 0; i < 3L * np; i++) fprintf(f, "%.9g ", pos[i]);
    fprintf(f, "\n");
    for (long i = 0; i < 3L * np; i++) fprintf(f, "%.9g ", vel ? vel[i] : 0.0);
    fprintf(f, "\n");
    for (long i = 0; i < np; i++) fprintf(f, "%.9g ", press ? press[i] : 0.0);
    fprintf(f, "\n");
    fclose(f);
    return MPS_OK;
}

// splitmix64 finaliser of (seed, counter): the jitter of coordinate a of
// particle i uses counter 3 i + a.
static inline double jitter_unit(unsigned long long seed, unsigned long long ctr)
{
    unsigned long long z = seed + (ctr + 1ull) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

extern "C" int mps_scene_dambreak(int dim, int fx, int fy, int fz, int nx, int ny, int nz,
                                  int layers_wall, int layers_dummy, double dp, double jitter,
                                  unsigned long long seed, int* np, int* nb2, int* nb1, double* pos)
{
    if ((dim != 2 && dim != 3) || fx < 0 || fz < 0 || nx <= 0 || nz <= 0 || layers_wall < 0 ||
        layers_dummy < 0 || !np || !nb2 || !nb1)
        return MPS_ERR_ARG;
    if (dim == 3 && (ny <= 0 || fy < 0)) return MPS_ERR_ARG;
    const int L = layers_wall + layers_dummy;
    const int y0 = dim == 3 ? -L : 0, y1 = dim == 3 ? ny + L : 1;
    long cnt[3] = {0, 0, 0};
    // pass 0 counts, pass 1 writes dummy, wall, fluid blocks in that order
    for (int pass = 0; pass < 2; pass++) {
        long off[3] = {cnt[2] + cnt[1], cnt[2], 0};   // fluid, wall, dummy start
        long k_out[3] = {0, 0, 0};
        if (pass == 1 && !pos) break;
        for (int k = -L; k < nz; k++)
            for (int j = y0; j < y1; j++)
                for (int i = -L; i < nx + L; i++) {
                    int di = i < 0 ? -i : (i >= nx ? i - nx + 1 : 0);
                    int dk = k < 0 ? -k : 0;
                    int dj = dim == 3 ? (j < 0 ? -j : (j >= ny ? j - ny + 1 : 0)) : 0;
                    int d = std::max(di, std::max(dj, dk));
                    int t;
                    if (d == 0) {
                        if (!(i < fx && k < fz && (dim == 2 || j < fy))) continue;
                        t = 0;
                    } else {
                        t = d <= layers_wall ? 1 : 2;
                    }
                    if (pass == 0) { cnt[t]++; continue; }
                    long idx = off[t] + k_out[t]++;
                    double q[3] = {i * dp, dim == 3 ? j * dp : 0.0, k * dp};
                    for (int a = 0; a < 3; a++) {
                        if (a == 1 && dim == 2) continue;
                        q[a] += jitter * dp * jitter_unit(seed, 3ull * (unsigned long long)idx + a);
                    }
                    pos[3 * idx + 0] = q[0]; pos[3 * idx + 1] = q[1]; pos[3 * idx + 2] = q[2];
                }
    }
    *np = (int)(cnt[0] + cnt[1] + cnt[2]);
    *nb2 = (int)cnt[2];
    *nb1 = (int)cnt[1];
    return MPS_OK;
}

// Wave tank (3-D, lattice units of dp): water x in [0, nx), y in [0, ny),
// z in [0, depth) in an open-top tank of inner nx x ny x nz.  The left wall is
// a piston: its layers (x < 0 over the inner y, z range) are flagged in
// `piston` and may travel `travel` lattice steps either way, so the floor and
// the side walls extend `travel` steps behind it.  Layers as in
// mps_scene_dambreak (wall nearest the water, dummies outside); particles come
// out as dummy, wall, fluid blocks.
extern "C" int mps_scene_wavetank(int nx, int ny, int nz, int depth, int travel, int layers_wall,
                                  int layers_dummy, double dp, double jitter, unsigned long long seed,
                                  int* np, int* nb2, int* nb1, double* pos, unsigned char* piston)
{
    if (nx <= 0 || ny <= 0 || nz <= 0 || depth < 0 || depth > nz || travel < 0 || layers_wall < 0 ||
        layers_dummy < 0 || !np || !nb2 || !nb1)
        return MPS_ERR_ARG;
    const int L = layers_wall + layers_dummy;
    long cnt[3] = {0, 0, 0};
    for (int pass = 0; pass < 2; pass++) {
        long off[3] = {cnt[2] + cnt[1], cnt[2], 0};   // fluid, wall, dummy start
        long k_out[3] = {0, 0, 0};
        if (pass == 1 && !pos) break;
        for (int k = -L; k < nz; k++)
            for (int j = -L; j < ny + L; j++)
                for (int i = -L - travel; i < nx + L; i++) {
                    const bool in_y = j >= 0 && j < ny, in_z = k >= 0;
                    const int dj = j < 0 ? -j : (j >= ny ? j - ny + 1 : 0);
                    const int dk = k < 0 ? -k : 0;
                    int d, t;
                    bool pist = false;
                    if (i >= 0 && i < nx && in_y && in_z) {          // inside the tank
                        if (k >= depth) continue;
                        d = 0;
                    } else if (i < 0 && in_y && in_z) {              // piston or the space behind it
                        if (i < -L) continue;
                        d = -i;
                        pist = true;
                    } else {                                         // floor, side walls, far wall
                        const int di = i >= nx ? i - nx + 1 : 0;
                        d = std::max(di, std::max(dj, dk));
                    }
                    t = d == 0 ? 0 : (d <= layers_wall ? 1 : 2);
                    if (pass == 0) { cnt[t]++; continue; }
                    const long idx = off[t] + k_out[t]++;
                    double q[3] = {i * dp, j * dp, k * dp};
                    for (int a = 0; a < 3; a++)
                        q[a] += jitter * dp * jitter_unit(seed, 3ull * (unsigned long long)idx + a);
                    pos[3 * idx + 0] = q[0]; pos[3 * idx + 1] = q[1]; pos[3 * idx + 2] = q[2];
                    if (piston) piston[idx] = pist ? 1 : 0;
                }
    }
    *np = (int)(cnt[0] + cnt[1] + cnt[2]);
    *nb2 = (int)cnt[2];
    *nb1 = (int)cnt[1];
    return MPS_OK;
}

// ============================================================== context ===

namespace {

float smallest_f(float a, bool strict)
{
    // smallest float t with sqrtf(t) >= a (strict: > a)
    float t = a * a;
    auto ok = [&](float v) { float s = sqrtf(v); return strict ? s > a : s >= a; };
    while (t > 0.0f && ok(t)) t = nextafterf(t, 0.0f);
    while (!ok(t)) t = nextafterf(t, INFINITY);
    return t;
}
double smallest_d(double a, bool strict)
{
    double t = a * a;
    auto ok = [&](double v) { double s = sqrt(v); return strict ? s > a : s >= a; };
    while (t > 0.0 && ok(t)) t = nextafter(t, 0.0);
    while (!ok(t)) t = nextafter(t, INFINITY);
    return t;
}

}  // namespace

static int dmalloc(mps_ctx* c, void** p, size_t bytes)
{
    HIPOK(hipMalloc(p, bytes ? bytes : 16));
    c->allocs.push_back(*p);
    return MPS_OK;
}

static void tic(mps_ctx* c, int k) { if (c->timing) hipEventRecord(c->ev[k][0], c->st); }
static void toc(mps_ctx* c, int k)
{
    if (!c->timing) return;
    hipEventRecord(c->ev[k][1], c->st);
    hipEventSynchronize(c->ev[k][1]);
    float t = 0.0f;
    hipEventElapsedTime(&t, c->ev[k][0], c->ev[k][1]);
    c->ms[k] += t;
}

// MPS::allocateCellList (MPS_CAL_expl_v1.0/MPS.cpp:456-531) in the state precision
template <typename Real>
static int grid_init(const mps_params& P, Real h, int np, const double* pos, Grid* g)
{
    Real l = (Real)pos[0], r = l, bk = (Real)pos[1], fr = bk, bo = (Real)pos[2], tp = bo;
    for (int i = 0; i < np; i++) {
        Real a = (Real)pos[3 * i], b = (Real)pos[3 * i + 1], cc = (Real)pos[3 * i + 2];
        l = std::min(l, a); r = std::max(r, a);
        bk = std::min(bk, b); fr = std::max(fr, b);
        bo = std::min(bo, cc); tp = std::max(tp, cc);
    }
    const Real md = (Real)P.max_dis;
    l -= md; r += md; bk -= md; fr += md; bo -= md; tp += md;
    if (P.dim == 2) { bk = (Real)0; fr = (Real)0; }
    const Real Lx = r - l, Ly = fr - bk, Lz = tp - bo;
    g->dx = (int)(unsigned)(Lx / h) + 1;
    g->dy = (int)(unsigned)(Ly / h) + 1;
    g->dz = (int)(unsigned)(Lz / h) + 1;
    g->left = l; g->back = bk; g->bottom = bo; g->h = h;
    g->sheet = g->dx * g->dy;
    g->ncells = g->sheet * g->dz;
    if (g->dx <= 2 || g->dz <= 2) return MPS_ERR_GRID;
    if (g->dy <= 2 && P.dim == 3) return MPS_ERR_GRID;
    return MPS_OK;
}

template <typename Real>
static Thresh<Real> make_thresh(const mps_params& P);
template <>
Thresh<float> make_thresh<float>(const mps_params& P)
{
    Thresh<float> t{};
    const float dp = (float)P.dp;
    t.re_n = (float)P.re_n_mult * dp;
    t.re_lap = (float)P.re_lap_mult * dp;
    t.col = (float)P.col_mult * dp;
    t.re_n2 = t.re_n * t.re_n;
    t.t_n = smallest_f(t.re_n, false);
    t.t_lap = smallest_f(t.re_lap, false);
    t.t_col = smallest_f(t.col, true);
    return t;
}
template <>
Thresh<double> make_thresh<double>(const mps_params& P)
{
    Thresh<double> t{};
    t.re_n = P.re_n_mult * P.dp;
    t.re_lap = P.re_lap_mult * P.dp;
    t.col = P.col_mult * P.dp;
    t.re_n2 = t.re_n * t.re_n;
    t.t_n = smallest_d(t.re_n, false);
    t.t_lap = smallest_d(t.re_lap, false);
    t.t_col = smallest_d(t.col, true);
    return t;
}

// Re-binning (MPS_GPU::divideCell + sort_*).  The first sort of a state the
// host supplied runs the radix sort; every later one starts from the previous
// sort's order and gathers each new cell from the 27 old cells around it.
// Both give the stable permutation of the current order by cell key.
template <typename Real>
static int divide_t(mps_ctx* c)
{
    StateView<Real> cur = c->view<Real>(c->kin, c->aux);
    StateView<Real> dst = c->view<Real>(1 - c->kin, 1 - c->aux);
    const bool gather = c->binned;
    launch_hash<Real>(c->g, c->nt, cur, c->k0, c->v0, &c->sc->nexcl, gather ? (uint32_t*)c->cs_next : nullptr,
                      c->st);
    uint32_t* vr;
    if (gather) {
        launch_gather_sort(c->g, c->nt, c->k0, c->cs, (uint32_t*)c->cs_next, c->cs_tiles, c->v1, c->st);
        vr = c->v1;
    } else {
        uint32_t* kr;
        launch_radix_sort(c->nt, c->key_bits, c->k0, c->v0, c->k1, c->v1, c->hist, &kr, &vr, c->st);
        launch_cell_bounds(kr, c->nt, c->g.ncells, c->cs_next, c->st);
    }
    std::swap(c->cs, c->cs_next);
    launch_reorder<Real>(c->nt, vr, cur, dst, c->inv, c->ntmp, c->cs_next, c->g.ncells + 1, c->st);
    launch_tail_copy<Real>(c->nt, c->cs + c->g.ncells, dst, cur, c->st);
    c->perm = vr;
    c->kin ^= 1;
    c->aux ^= 1;
    c->binned = true;
    return hipGetLastError() == hipSuccess ? MPS_OK : MPS_ERR_HIP;
}

static int divide(mps_ctx* c) { return c->prec ? divide_t<double>(c) : divide_t<float>(c); }

template <typename Real>
static const Thresh<Real>& thr(mps_ctx* c);
template <> const Thresh<float>& thr<float>(mps_ctx* c) { return c->th.f; }
template <> const Thresh<double>& thr<double>(mps_ctx* c) { return c->th.d; }

enum Stage { S_VISC, S_COL, S_PND, S_DUMMY, S_GRAD, S_PND_ONLY };

template <typename Real>
static int run_stage_t(mps_ctx* c, Stage s)
{
    StateView<Real> cur = c->view<Real>(c->kin, c->aux);
    StateView<Real> alt = c->view<Real>(1 - c->kin, c->aux);
    switch (s) {
    case S_VISC:
        tic(c, T_VISC);
        launch_visc<Real>(c->g, c->cs, cur, alt, thr<Real>(c), c->kc, c->sc, c->st);
        c->kin ^= 1;
        toc(c, T_VISC);
        break;
    case S_COL:
        tic(c, T_COL);
        // the hit counter is cleared by the dt kernel at the start of every step
        // (and by mps_stage_collision when the stage runs on its own)
        launch_collision<Real>(c->g, c->cs, cur, alt, thr<Real>(c), c->kc, c->sc, c->st);
        c->kin ^= 1;
        toc(c, T_COL);
        break;
    case S_PND:
        tic(c, T_PND);
        launch_pnd<Real>(c->g, c->cs, cur, thr<Real>(c), c->kc, c->sc, 0, 1, c->st);
        toc(c, T_PND);
        break;
    case S_DUMMY:
        tic(c, T_DUMMY);
        launch_dummy_p<Real>(c->nt, cur, c->st);
        toc(c, T_DUMMY);
        break;
    case S_PND_ONLY:
        tic(c, T_PND);
        launch_pnd<Real>(c->g, c->cs, cur, thr<Real>(c), c->kc, c->sc, 0, 0, c->st);
        toc(c, T_PND);
        break;
    case S_GRAD:
        tic(c, T_GRAD);
        launch_gradp<Real>(c->g, c->cs, cur, alt, thr<Real>(c), c->kc, c->sc, c->st);
        c->kin ^= 1;
        toc(c, T_GRAD);
        break;
    }
    return hipGetLastError() == hipSuccess ? MPS_OK : MPS_ERR_HIP;
}

static int run_stage(mps_ctx* c, Stage s) { return c->prec ? run_stage_t<double>(c, s) : run_stage_t<float>(c, s); }

static int do_sort(mps_ctx* c)
{
    tic(c, T_SORT);
    int rc = divide(c);
    toc(c, T_SORT);
    return rc;
}

int mps_ctx_divide(mps_ctx* c) { return do_sort(c); }
int mps_ctx_stage(mps_ctx* c, int stage) { return run_stage(c, (Stage)stage); }

// One step.  rebin_every_move = 0: MPS_CAL_/MPS_GPU.cpp:611-714 (re-bin, visc,
// collision, n, p, grad p); 1: MPS_CAL_expl_v1.0/MPS_GPU_EXPL.cpp:85-147
// (visc, re-bin, collision, re-bin, n, p, grad p, re-bin).  dt comes first in
// both (MPS_GPU_EXPL.cpp:95).
static int step_once(mps_ctx* c)
{
    int rc;
    tic(c, T_DT);
    launch_dt(c->kc, c->sc, c->st);
    toc(c, T_DT);
    if ((rc = mps_ctx_motion(c))) return rc;
    const bool every = c->P.rebin_every_move != 0;
    if (!every && (rc = do_sort(c))) return rc;
    if ((rc = run_stage(c, S_VISC))) return rc;
    if (every && (rc = do_sort(c))) return rc;
    if ((rc = run_stage(c, S_COL))) return rc;
    if (every && (rc = do_sort(c))) return rc;
    if ((rc = run_stage(c, S_PND))) return rc;
    if ((rc = run_stage(c, S_DUMMY))) return rc;
    if ((rc = run_stage(c, S_GRAD))) return rc;
    if (every && (rc = do_sort(c))) return rc;
    c->steps++;
    return MPS_OK;
}

static int check_device()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return MPS_ERR_NODEV;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return MPS_ERR_NODEV;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return MPS_ERR_NODEV;
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        fprintf(stderr, "mps_amd: device %s is not gfx950\n", prop.gcnArchName);
        return MPS_ERR_NODEV;
    }
    return MPS_OK;
}

template <typename Real>
static int upload_t(mps_ctx* c, int k, int a, const double* pos, const double* vel, const double* press)
{
    const int N = c->nt;
    // host-supplied positions can be anywhere: the next re-binning sorts from scratch
    if (pos) c->binned = false;
    std::vector<Real> h((size_t)N);
    for (int q = 0; q < 3; q++) {
        if (pos) {
            for (int i = 0; i < N; i++) h[i] = (Real)pos[3 * i + q];
            HIPOK(hipMemcpy(c->f[k][q], h.data(), sizeof(Real) * N, hipMemcpyHostToDevice));
        }
        if (vel) {
            for (int i = 0; i < N; i++) h[i] = (Real)vel[3 * i + q];
            HIPOK(hipMemcpy(c->f[k][3 + q], h.data(), sizeof(Real) * N, hipMemcpyHostToDevice));
        }
    }
    if (press) {
        for (int i = 0; i < N; i++) h[i] = (Real)press[i];
        HIPOK(hipMemcpy(c->f[a][6], h.data(), sizeof(Real) * N, hipMemcpyHostToDevice));
    }
    return MPS_OK;
}

template <typename Real>
static int calonce_t(mps_ctx* c)
{
    launch_calonce<Real>(c->g, c->cs, c->nt, c->view<Real>(c->kin, c->aux), thr<Real>(c), c->kc, c->sc,
                         c->st);
    return hipGetLastError() == hipSuccess ? MPS_OK : MPS_ERR_HIP;
}

extern "C" int mps_create(const mps_params* p, int np, int nb2, int nb1, const double* pos,
                          const double* vel, const double* press, mps_ctx** out)
{
    if (!p || !out || np <= 0 || nb2 < 0 || nb1 < 0 || nb2 + nb1 > np || !pos) return MPS_ERR_ARG;
    if (p->dim != 2 && p->dim != 3) return MPS_ERR_ARG;
    *out = nullptr;
    int rc = check_device();
    if (rc) return rc;
    mps_ctx* c = new mps_ctx();
    c->P = *p;
    c->prec = p->precision ? 1 : 0;
    c->nt = np; c->nb2 = nb2; c->nb1 = nb1;
    c->th.f = make_thresh<float>(*p);
    c->th.d = make_thresh<double>(*p);
    if (c->prec) rc = grid_init<double>(*p, std::max(c->th.d.re_n, c->th.d.re_lap), np, pos, &c->g);
    else rc = grid_init<float>(*p, std::max(c->th.f.re_n, c->th.f.re_lap), np, pos, &c->g);
    if (rc) { delete c; return rc; }
    c->key_bits = 0;
    while ((1u << c->key_bits) <= (unsigned)c->g.ncells) c->key_bits++;

    StepConst& k = c->kc;
    k.rho = p->rho; k.nu = p->nu; k.cs = p->cs; k.cfl = p->cfl; k.dp = p->dp; k.dt_max = p->dt_max;
    k.dt_min = c->prec ? p->dt_min_mult * p->dt_max : (double)((float)p->dt_min_mult * (float)p->dt_max);
    k.col_rate = p->col_rate;
    k.g[0] = p->g[0]; k.g[1] = p->g[1]; k.g[2] = p->g[2];
    k.dim = p->dim;
    k.gradp_skip_dummy = p->gradp_skip_dummy;
    double fl2 = 0.0, st2 = 0.0;
    for (int i = 0; i < np; i++) {
        double a = vel ? vel[3 * i] * vel[3 * i] + vel[3 * i + 1] * vel[3 * i + 1] + vel[3 * i + 2] * vel[3 * i + 2] : 0.0;
        if (i >= nb2 + nb1) fl2 = std::max(fl2, a); else st2 = std::max(st2, a);
    }
    k.umax2_static = st2;

    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) { delete c; return MPS_ERR_HIP; }
    const size_t N = (size_t)np, rs = c->rsize();
    rc = 0;
    for (int s = 0; s < 2 && !rc; s++) {
        for (int q = 0; q < 7 && !rc; q++) rc = dmalloc(c, &c->f[s][q], rs * N);
        if (!rc) rc = dmalloc(c, (void**)&c->n[s], 8 * N);
        if (!rc) rc = dmalloc(c, (void**)&c->type[s], 4 * N);
        if (!rc) rc = dmalloc(c, (void**)&c->id[s], 4 * N);
        if (!rc) rc = dmalloc(c, (void**)&c->normal[s], 4 * N);
    }
    if (!rc) rc = dmalloc(c, (void**)&c->k0, 4 * N);
    if (!rc) rc = dmalloc(c, (void**)&c->v0, 4 * N);
    if (!rc) rc = dmalloc(c, (void**)&c->k1, 4 * N);
    if (!rc) rc = dmalloc(c, (void**)&c->v1, 4 * N);
    if (!rc) rc = dmalloc(c, (void**)&c->hist, 4 * (size_t)sort_scratch_words(np));
    if (!rc) rc = dmalloc(c, (void**)&c->inv, 4 * N);
    if (!rc) rc = dmalloc(c, (void**)&c->ntmp, 4 * N);
    if (!rc) rc = dmalloc(c, (void**)&c->cs, 4 * ((size_t)c->g.ncells + 1));
    if (!rc) rc = dmalloc(c, (void**)&c->cs_next, 4 * ((size_t)c->g.ncells + 1));
    if (!rc) rc = dmalloc(c, (void**)&c->cs_tiles, 4 * (size_t)scan_tile_words(c->g.ncells + 1));
    if (!rc) rc = dmalloc(c, (void**)&c->sc, sizeof(DevScalars));
    if (rc) { mps_destroy(c); return rc; }
    // every remaining set-up call is checked: a context is returned whole or not at all
    bool good = true;
    auto ok = [&good](hipError_t e) { good = good && e == hipSuccess; };
    for (int t = 0; t < T_N; t++) {
        ok(hipEventCreate(&c->ev[t][0]));
        ok(hipEventCreate(&c->ev[t][1]));
    }

    // state: types by index range (MPS.cpp:263-268), id = i, no normals yet
    std::vector<int> ht(N), hid(N), hn(N, -1);
    for (int i = 0; i < np; i++) { ht[i] = i < nb2 ? 2 : (i < nb2 + nb1 ? 1 : 0); hid[i] = i; }
    if ((rc = c->prec ? upload_t<double>(c, 0, 0, pos, vel, press) : upload_t<float>(c, 0, 0, pos, vel, press))) {
        mps_destroy(c);
        return rc;
    }
    if (!vel)
        for (int q = 3; q < 6; q++) ok(hipMemset(c->f[0][q], 0, rs * N));
    if (!press) ok(hipMemset(c->f[0][6], 0, rs * N));
    ok(hipMemcpy(c->type[0], ht.data(), 4 * N, hipMemcpyHostToDevice));
    ok(hipMemcpy(c->id[0], hid.data(), 4 * N, hipMemcpyHostToDevice));
    ok(hipMemcpy(c->normal[0], hn.data(), 4 * N, hipMemcpyHostToDevice));
    ok(hipMemset(c->n[0], 0, 8 * N));
    ok(hipMemset(c->cs_next, 0, 4 * ((size_t)c->g.ncells + 1)));
    DevScalars s0{};
    s0.dt = p->dt_max;
    s0.maxu2_bits = 0;
    {
        float f2 = (float)fl2;
        memcpy(&s0.maxu2_bits, &f2, 4);
    }
    s0.argmax = 0x7fffffff;
    ok(hipMemcpy(c->sc, &s0, sizeof s0, hipMemcpyHostToDevice));
    if (!good) {
        fprintf(stderr, "mps_amd: mps_create: device set-up failed\n");
        mps_destroy(c);
        return MPS_ERR_HIP;
    }

    // MPS_GPU::Initial order (MPS_GPU.cpp:100-106): divideCell, makeLink, calOnce
    if ((rc = divide(c)) || (rc = c->prec ? calonce_t<double>(c) : calonce_t<float>(c))) {
        mps_destroy(c);
        return rc;
    }
    if (hipStreamSynchronize(c->st) != hipSuccess) { mps_destroy(c); return MPS_ERR_HIP; }
    *out = c;
    return MPS_OK;
}

extern "C" int mps_destroy(mps_ctx* c)
{
    if (!c) return MPS_OK;
    if (c->st) hipStreamSynchronize(c->st);
    mps_ppe_free(c);
    mps_mlp_free(c);
    for (void* p : c->allocs) hipFree(p);
    for (int t = 0; t < T_N; t++)
        for (int e = 0; e < 2; e++)
            if (c->ev[t][e]) hipEventDestroy(c->ev[t][e]);
    if (c->st) hipStreamDestroy(c->st);
    delete c;
    return MPS_OK;
}

extern "C" int mps_sync(mps_ctx* c)
{
    if (!c) return MPS_ERR_ARG;
    HIPOK(hipStreamSynchronize(c->st));
    int err = 0;
    HIPOK(hipMemcpy(&err, &c->sc->err, sizeof(int), hipMemcpyDeviceToHost));
    return err ? MPS_ERR_DT : MPS_OK;
}

extern "C" int mps_step_async(mps_ctx* c, int nsteps)
{
    if (!c || nsteps < 0) return MPS_ERR_ARG;
    for (int s = 0; s < nsteps; s++) {
        if (c->timing) hipEventRecord(c->ev[T_STEP][0], c->st);
        int rc = step_once(c);
        if (rc) return rc;
        if (c->timing) {
            hipEventRecord(c->ev[T_STEP][1], c->st);
            hipEventSynchronize(c->ev[T_STEP][1]);
            float t = 0.0f;
            hipEventElapsedTime(&t, c->ev[T_STEP][0], c->ev[T_STEP][1]);
            c->ms[T_STEP] += t;
        }
    }
    return MPS_OK;
}

extern "C" int mps_step(mps_ctx* c, int nsteps)
{
    int rc = mps_step_async(c, nsteps);
    if (rc) return rc;
    return mps_sync(c);
}

extern "C" int mps_get_np(mps_ctx* c, int* np_total, int* np_active)
{
    if (!c) return MPS_ERR_ARG;
    if (np_total) *np_total = c->nt;
    if (np_active) {
        HIPOK(hipStreamSynchronize(c->st));
        HIPOK(hipMemcpy(np_active, c->cs + c->g.ncells, sizeof(int), hipMemcpyDeviceToHost));
    }
    return MPS_OK;
}

template <typename Real>
static int get_state_t(mps_ctx* c, int order, double* pos, double* vel, double* press, double* n,
                       int* id, int* type)
{
    const int N = c->nt;
    std::vector<int> hid((size_t)N);
    HIPOK(hipMemcpy(hid.data(), c->id[c->aux], 4 * (size_t)N, hipMemcpyDeviceToHost));
    auto dst = [&](int i) { return order ? hid[i] : i; };
    std::vector<Real> h((size_t)N);
    for (int q = 0; q < 7; q++) {
        double* o = q < 3 ? pos : (q < 6 ? vel : press);
        if (!o) continue;
        HIPOK(hipMemcpy(h.data(), c->f[q < 6 ? c->kin : c->aux][q], sizeof(Real) * N, hipMemcpyDeviceToHost));
        if (q < 6) for (int i = 0; i < N; i++) o[3 * dst(i) + (q % 3)] = (double)h[i];
        else for (int i = 0; i < N; i++) o[dst(i)] = (double)h[i];
    }
    if (n) {
        std::vector<double> hn((size_t)N);
        HIPOK(hipMemcpy(hn.data(), c->n[c->aux], 8 * (size_t)N, hipMemcpyDeviceToHost));
        for (int i = 0; i < N; i++) n[dst(i)] = hn[i];
    }
    if (type) {
        std::vector<int> ht((size_t)N);
        HIPOK(hipMemcpy(ht.data(), c->type[c->aux], 4 * (size_t)N, hipMemcpyDeviceToHost));
        for (int i = 0; i < N; i++) type[dst(i)] = ht[i];
    }
    if (id) for (int i = 0; i < N; i++) id[dst(i)] = hid[i];
    return MPS_OK;
}

extern "C" int mps_get_state(mps_ctx* c, int order, double* pos, double* vel, double* press, double* n,
                             int* id, int* type)
{
    if (!c || (order != 0 && order != 1)) return MPS_ERR_ARG;
    HIPOK(hipStreamSynchronize(c->st));
    return c->prec ? get_state_t<double>(c, order, pos, vel, press, n, id, type)
                   : get_state_t<float>(c, order, pos, vel, press, n, id, type);
}

extern "C" int mps_set_state(mps_ctx* c, const double* pos, const double* vel, const double* press)
{
    if (!c) return MPS_ERR_ARG;
    HIPOK(hipStreamSynchronize(c->st));
    return c->prec ? upload_t<double>(c, c->kin, c->aux, pos, vel, press)
                   : upload_t<float>(c, c->kin, c->aux, pos, vel, press);
}

extern "C" int mps_get_scalars(mps_ctx* c, double* out)
{
    if (!c || !out) return MPS_ERR_ARG;
    HIPOK(hipStreamSynchronize(c->st));
    DevScalars s;
    HIPOK(hipMemcpy(&s, c->sc, sizeof s, hipMemcpyDeviceToHost));
    out[0] = s.n0; out[1] = s.lambda; out[2] = s.coef; out[3] = s.dt; out[4] = s.time;
    out[5] = c->g.ncells; out[6] = s.ncol; out[7] = s.nexcl; out[8] = c->g.h; out[9] = (double)c->steps;
    return MPS_OK;
}

extern "C" int mps_get_grid(mps_ctx* c, double* geom, int* dims)
{
    if (!c) return MPS_ERR_ARG;
    if (geom) { geom[0] = c->g.left; geom[1] = c->g.back; geom[2] = c->g.bottom; geom[3] = c->g.h; }
    if (dims) { dims[0] = c->g.dx; dims[1] = c->g.dy; dims[2] = c->g.dz; dims[3] = c->g.ncells; }
    return MPS_OK;
}

// "%d %d %lf %lf %lf %lf %lf %lf %lf " per active particle in the current order
extern "C" int mps_write_out(mps_ctx* c, const char* path)
{
    if (!c || !path) return MPS_ERR_ARG;
    int na = 0;
    int rc = mps_get_np(c, nullptr, &na);
    if (rc) return rc;
    const size_t N = (size_t)c->nt;
    std::vector<double> pos(3 * N), vel(3 * N), pr(N);
    std::vector<int> id(N), ty(N);
    if ((rc = mps_get_state(c, 0, pos.data(), vel.data(), pr.data(), nullptr, id.data(), ty.data()))) return rc;
    FILE* f = fopen(path, "wt");
    if (!f) return MPS_ERR_IO;
    for (int i = 0; i < na; i++)
        fprintf(f, "%d %d %lf %lf %lf %lf %lf %lf %lf ", id[i], ty[i], pos[3 * i], pos[3 * i + 1],
                pos[3 * i + 2], vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], pr[i]);
    fclose(f);
    return MPS_OK;
}

extern "C" int mps_write_vtk(mps_ctx* c, const char* path)
{
    if (!c || !path) return MPS_ERR_ARG;
    // every particle of GEO.dat, in id order (excluded ones at their last position)
    const int na = c->nt;
    const size_t N = (size_t)c->nt;
    std::vector<double> pos(3 * N), vel(3 * N), pr(N);
    std::vector<int> id(N);
    int rc;
    if ((rc = mps_get_state(c, 1, pos.data(), vel.data(), pr.data(), nullptr, id.data(), nullptr))) return rc;
    FILE* f = fopen(path, "wt");
    if (!f) return MPS_ERR_IO;
    fprintf(f, "# vtk DataFile Version 3.0\nunstructured grid\nASCII\nDATASET UNSTRUCTURED_GRID\n");
    fprintf(f, "POINTS %d float\n", na);
    for (int i = 0; i < na; i++) fprintf(f, "%lf %lf %lf ", pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    fprintf(f, "\nCELLS %d %d\n", na, 2 * na);
    for (int i = 0; i < na; i++) fprintf(f, "%d %d ", 1, i);
    fprintf(f, "\nCELL_TYPES %d\n", na);
    for (int i = 0; i < na; i++) fprintf(f, "1 ");
    fprintf(f, "\nPOINT_DATA %d\nSCALARS id int 1\nLOOKUP_TABLE default\n", na);
    for (int i = 0; i < na; i++) fprintf(f, "%d ", id[i]);
    fprintf(f, "\nSCALARS pressure float 1\nLOOKUP_TABLE default\n");
    for (int i = 0; i < na; i++) fprintf(f, "%lf ", pr[i]);
    fprintf(f, "\nVECTORS velocity float\n");
    for (int i = 0; i < na; i++) fprintf(f, "%lf %lf %lf ", vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
    fprintf(f, "\n");
    const bool ok = !ferror(f);
    fclose(f);
    return ok ? MPS_OK : MPS_ERR_IO;
}

int mps_ctx_motion(mps_ctx* c)
{
    if (!c->mov || c->kc.mov_amp == 0.0) return MPS_OK;
    if (c->prec)
        launch_motion<double>(c->nt, c->view<double>(c->kin, c->aux), c->mov, c->xref, c->mov_axis,
                              c->kc.mov_amp, c->kc.mov_omega, c->sc, c->st);
    else
        launch_motion<float>(c->nt, c->view<float>(c->kin, c->aux), c->mov, c->xref, c->mov_axis,
                             c->kc.mov_amp, c->kc.mov_omega, c->sc, c->st);
    HIPOK(hipGetLastError());
    return MPS_OK;
}

extern "C" int mps_set_motion(mps_ctx* c, const unsigned char* moving, int axis, double amplitude, double freq)
{
    if (!c || axis < 0 || axis > 2 || !std::isfinite(amplitude) || !std::isfinite(freq)) return MPS_ERR_ARG;
    if (!moving || amplitude == 0.0) {
        c->kc.mov_amp = 0.0;
        return MPS_OK;
    }
    const size_t N = (size_t)c->nt;
    int rc;
    if (!c->mov) {
        if ((rc = dmalloc(c, (void**)&c->mov, N))) return rc;
        if ((rc = dmalloc(c, (void**)&c->xref, 8 * N))) return rc;
    }
    HIPOK(hipMemcpyAsync(c->mov, moving, N, hipMemcpyHostToDevice, c->st));
    c->mov_axis = axis;
    c->kc.mov_amp = amplitude;
    c->kc.mov_omega = 2.0 * M_PI * freq;
    if (c->prec)
        launch_motion_ref<double>(c->nt, c->view<double>(c->kin, c->aux), c->mov, c->xref, axis, amplitude,
                                  c->kc.mov_omega, c->sc, c->st);
    else
        launch_motion_ref<float>(c->nt, c->view<float>(c->kin, c->aux), c->mov, c->xref, axis, amplitude,
                                 c->kc.mov_omega, c->sc, c->st);
    HIPOK(hipGetLastError());
    HIPOK(hipStreamSynchronize(c->st));
    return MPS_OK;
}

extern "C" int mps_stage_sort(mps_ctx* c, int* perm, int* cell_start)
{
    if (!c) return MPS_ERR_ARG;
    int rc = do_sort(c);
    if (rc) return rc;
    HIPOK(hipStreamSynchronize(c->st));
    if (perm) HIPOK(hipMemcpy(perm, c->perm, 4 * (size_t)c->nt, hipMemcpyDeviceToHost));
    if (cell_start) HIPOK(hipMemcpy(cell_start, c->cs, 4 * ((size_t)c->g.ncells + 1), hipMemcpyDeviceToHost));
    return MPS_OK;
}

static int stage_sync(mps_ctx* c, Stage s)
{
    if (!c) return MPS_ERR_ARG;
    int rc = run_stage(c, s);
    if (rc) return rc;
    HIPOK(hipStreamSynchronize(c->st));
    return MPS_OK;
}

extern "C" int mps_stage_visc(mps_ctx* c) { return stage_sync(c, S_VISC); }
extern "C" int mps_stage_collision(mps_ctx* c)
{
    if (!c) return MPS_ERR_ARG;
    HIPOK(hipMemsetAsync(&c->sc->ncol, 0, sizeof(int), c->st));
    return stage_sync(c, S_COL);
}
extern "C" int mps_stage_pnd_pres(mps_ctx* c)
{
    int rc = stage_sync(c, S_PND);
    return rc ? rc : stage_sync(c, S_DUMMY);
}
extern "C" int mps_stage_gradp(mps_ctx* c) { return stage_sync(c, S_GRAD); }

extern "C" int mps_stage_dt(mps_ctx* c)
{
    if (!c) return MPS_ERR_ARG;
    launch_dt(c->kc, c->sc, c->st);
    return mps_sync(c);
}

extern "C" int mps_set_dt(mps_ctx* c, double dt)
{
    if (!c || !(dt > 0.0)) return MPS_ERR_ARG;
    HIPOK(hipStreamSynchronize(c->st));
    HIPOK(hipMemcpy(&c->sc->dt, &dt, sizeof(double), hipMemcpyHostToDevice));
    return MPS_OK;
}

template <typename Real>
static int nlist_t(mps_ctx* c, int which, int* row_ptr, int* col, long cap, long* nnz)
{
    int na = 0;
    int rc = mps_get_np(c, nullptr, &na);
    if (rc) return rc;
    int *drp = nullptr, *dcol = nullptr;
    HIPOK(hipMalloc(&drp, 4 * ((size_t)na + 1)));
    HIPOK(hipMemset(drp, 0, 4 * ((size_t)na + 1)));
    StateView<Real> s = c->view<Real>(c->kin, c->aux);
    launch_nlist<Real>(c->g, c->cs, na, s, thr<Real>(c), which, drp, nullptr, 0, c->st);
    HIPOK(hipStreamSynchronize(c->st));
    std::vector<int> rp((size_t)na + 1);
    HIPOK(hipMemcpy(rp.data(), drp, 4 * ((size_t)na + 1), hipMemcpyDeviceToHost));
    long acc = 0;
    for (int i = 0; i < na; i++) { long t = rp[i + 1]; rp[i] = (int)acc; acc += t; }
    rp[na] = (int)acc;
    *nnz = acc;
    if (row_ptr) memcpy(row_ptr, rp.data(), 4 * ((size_t)na + 1));
    if (acc > cap || !col) { hipFree(drp); return acc > cap ? MPS_ERR_CAP : MPS_OK; }
    HIPOK(hipMemcpy(drp, rp.data(), 4 * ((size_t)na + 1), hipMemcpyHostToDevice));
    HIPOK(hipMalloc(&dcol, 4 * (size_t)std::max(acc, 1L)));
    launch_nlist<Real>(c->g, c->cs, na, s, thr<Real>(c), which, drp, dcol, 1, c->st);
    HIPOK(hipStreamSynchronize(c->st));
    HIPOK(hipMemcpy(col, dcol, 4 * (size_t)acc, hipMemcpyDeviceToHost));
    hipFree(drp);
    hipFree(dcol);
    return MPS_OK;
}

extern "C" int mps_neighbor_list(mps_ctx* c, int which, int* row_ptr, int* col, long cap, long* nnz)
{
    if (!c || !nnz || (which != 0 && which != 1)) return MPS_ERR_ARG;
    return c->prec ? nlist_t<double>(c, which, row_ptr, col, cap, nnz)
                   : nlist_t<float>(c, which, row_ptr, col, cap, nnz);
}

extern "C" int mps_stage_times(mps_ctx* c, double* ms, int reset)
{
    if (!c) return MPS_ERR_ARG;
    if (ms) for (int t = 0; t < T_N; t++) ms[t] = c->ms[t];
    if (reset) for (int t = 0; t < T_N; t++) c->ms[t] = 0.0;
    return MPS_OK;
}

extern "C" int mps_enable_timing(mps_ctx* c, int on)
{
    if (!c) return MPS_ERR_ARG;
    c->timing = on != 0;
    return MPS_OK;
}
