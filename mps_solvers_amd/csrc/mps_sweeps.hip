// mps_sweeps.hip -- the neighbour sweeps of the MPS step on gfx950.
//
// One 64-lane wavefront per cell (in chunks of 64 particles).  For each chunk
// the wave takes the bounding box of the particles the operator updates,
// streams the 27-cell stencil (9 contiguous rows of the cell-sorted arrays)
// with coalesced loads, keeps only candidates closer to the box than the
// operator's radius, compacts them into LDS with a ballot prefix, and then
// broadcasts each staged candidate to all 64 lanes.  Every particle's sum
// therefore sees exactly the reference's within-cutoff neighbour set.
//
// Operators (reference semantics, two-pass as in MPS_CAL_expl_v1.0):
//   PND + explicit pressure  MPS_GPU::cal_n (MPS_GPU.cpp:830-887) + calPres_expl
//                            (MPS_GPU_EXPL.cpp:199-238): n in FP64, p = max(0, rho c^2 (n-n0)/n0)
//   viscosity + gravity      calVisc_expl (MPS_GPU_EXPL.cpp:152-194), d3_LapVel (MPS_GPU.cpp:683-717)
//   collision                MPS_GPU::collision (MPS_GPU.cpp:892-986)
//   pressure gradient        MPS_GPU::calDash (MPS_GPU.cpp:782-825), d3_GradPres (MPS_GPU.cpp:571-638)
#include <float.h>

#include "mps_device.hpp"
#include "mps_kernels.hpp"
#include "mps_tile.hpp"

namespace mps {

// Round bookkeeping shared by the sweep kernels: lane -> (particle, sub-lane).
struct Lane {
    int lane, wave, sub;
    __device__ Lane() : lane(threadIdx.x & 63), wave(threadIdx.x >> 6), sub(threadIdx.x & 7) {}
    __device__ int particle(int c0) const { return c0 + wave * 8 + (lane >> 3); }
};

// ------------------------------------------------------------------ PND ---
template <typename Real>
struct OpPnd {
    const Real *x, *y, *z;
    Real tn;     // rr < tn  <=>  sqrt(rr) < R0
    double re;
    int i, sub;
    bool active;
    Real xi, yi, zi;
    double acc;

    __device__ void load_cand(int j, Cand<Real>& c) { c.x = x[j]; c.y = y[j]; c.z = z[j]; c.j = j; }
    __device__ void process(const Cand<Real>* S, const uint16_t* L, int len)
    {
        for (int k = sub; k < len; k += TK) {
            const Cand<Real>& c = S[L[k]];
            const Real cx = c.x, cy = c.y, cz = c.z;
            const int cj = c.j;
            const Real rr = rr3(cx - xi, cy - yi, cz - zi);
            if (active && rr < tn && cj != i) {
                // exact differences of the stored values (FP64 holds any float difference)
                const double ddx = (double)cx - (double)xi, ddy = (double)cy - (double)yi,
                             ddz = (double)cz - (double)zi;
                const double r2 = __builtin_fma(ddx, ddx, __builtin_fma(ddy, ddy, ddz * ddz));
                acc += __builtin_fma(re, rsq(r2), -1.0);
            }
        }
    }
};

template <typename Real>
__global__ void __launch_bounds__(TB) k_pnd(Grid g, const int* __restrict__ cs, StateView<Real> s,
                                            Thresh<Real> th, const DevScalars* __restrict__ sc,
                                            double one_over_alpha, int all_types, int with_pressure)
{
    __shared__ TileShared<Real> sh;
    const int cell = xcd_cell(g.ncells);
    if (cell >= g.ncells) return;
    const int s0 = cs[cell], s1 = cs[cell + 1];
    const double n0 = with_pressure ? sc->n0 : 1.0;
    const Lane ln;
    OpPnd<Real> op;
    op.x = s.x; op.y = s.y; op.z = s.z;
    op.re = (double)th.re_n; op.tn = th.t_n; op.sub = ln.sub;
    const Real R2 = filt_r2(th.re_n);
    for (int c0 = s0; c0 < s1; c0 += TW * 8) {
        const int i = ln.particle(c0);
        const bool in = i < s1;
        op.i = i;
        op.acc = 0.0;
        op.active = false;
        op.xi = op.yi = op.zi = (Real)0;
        if (in) {
            op.xi = s.x[i]; op.yi = s.y[i]; op.zi = s.z[i];
            op.active = all_types || s.type[i] != 2;
        }
        Real wlo[3], whi[3], blo[3], bhi[3];
        bool tact;
        if (!tile_boxes(sh, op.active, op.xi, op.yi, op.zi, wlo, whi, blo, bhi, tact)) continue;
        tile_stage(g, cs, cell, sh, blo, bhi, wlo, whi, R2, tact, op);
        const double acc = sub_sum(op.acc);
        if (op.active && ln.sub == 0) {
            s.n[i] = acc;
            if (with_pressure) {
                const double t = one_over_alpha * (acc - n0) / n0;
                s.p[i] = (Real)(t > 0.0 ? t : 0.0);
            }
        }
    }
}

// dummy pressure p_i = p_normal(i) (MPS_GPU_EXPL.cpp:228-236); no partner -> 0
template <typename Real>
__global__ void __launch_bounds__(256) k_dummy_p(int n, StateView<Real> s)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || s.type[i] != 2) return;
    const int m = s.normal[i];
    s.p[i] = m >= 0 ? s.p[m] : (Real)0;
}

// ------------------------------------------------------------ viscosity ---
template <typename Real>
struct OpVisc {
    const Real *x, *y, *z, *u, *v, *w;
    Real re, tl;
    int i, sub;
    bool active;
    Real xi, yi, zi, ui, vi, wi;
    Real lx, ly, lz;

    __device__ void load_cand(int j, Cand<Real>& c)
    {
        c.x = x[j]; c.y = y[j]; c.z = z[j]; c.j = j;
        c.a = u[j]; c.b = v[j]; c.c = w[j];
    }
    __device__ void process(const Cand<Real>* S, const uint16_t* L, int len)
    {
        for (int k = sub; k < len; k += TK) {
            const Cand<Real> c = S[L[k]];
            const Real dx = c.x - xi, dy = c.y - yi, dz = c.z - zi;
            const Real rr = rr3(dx, dy, dz);
            const bool inside = rr < tl && c.j != i;
            const Real wt = inside ? fmar(re, rsq(rr), (Real)-1) : (Real)0;
            lx = fmar(wt, c.a - ui, lx);
            ly = fmar(wt, c.b - vi, ly);
            lz = fmar(wt, c.c - wi, lz);
        }
    }
};

template <typename Real>
__global__ void __launch_bounds__(TB) k_visc(Grid g, const int* __restrict__ cs, StateView<Real> a,
                                             StateView<Real> b, Thresh<Real> th, StepConst kc,
                                             const DevScalars* __restrict__ sc)
{
    __shared__ TileShared<Real> sh;
    const int cell = xcd_cell(g.ncells);
    if (cell >= g.ncells) return;
    const int s0 = cs[cell], s1 = cs[cell + 1];
    const double dt = sc->dt, coef = sc->coef;
    const Lane ln;
    OpVisc<Real> op;
    op.x = a.x; op.y = a.y; op.z = a.z; op.u = a.u; op.v = a.v; op.w = a.w;
    op.re = th.re_lap; op.tl = th.t_lap; op.sub = ln.sub;
    const Real R2 = filt_r2(th.re_lap);
    for (int c0 = s0; c0 < s1; c0 += TW * 8) {
        const int i = ln.particle(c0);
        const bool in = i < s1;
        op.i = i;
        op.active = false;
        op.lx = op.ly = op.lz = (Real)0;
        op.xi = op.yi = op.zi = op.ui = op.vi = op.wi = (Real)0;
        if (in) {
            op.xi = a.x[i]; op.yi = a.y[i]; op.zi = a.z[i];
            op.ui = a.u[i]; op.vi = a.v[i]; op.wi = a.w[i];
            op.active = a.type[i] == 0;
        }
        Real wlo[3], whi[3], blo[3], bhi[3];
        bool tact;
        if (tile_boxes(sh, op.active, op.xi, op.yi, op.zi, wlo, whi, blo, bhi, tact))
            tile_stage(g, cs, cell, sh, blo, bhi, wlo, whi, R2, tact, op);
        const Real lx = sub_sum(op.lx), ly = sub_sum(op.ly), lz = sub_sum(op.lz);
        if (in && ln.sub == 0) {
            Real xo = op.xi, yo = op.yi, zo = op.zi, uo = op.ui, vo = op.vi, wo = op.wi;
            if (op.active) {
                const Real dtr = (Real)dt, nc = (Real)(kc.nu * coef);
                uo = op.ui + dtr * (nc * lx + (Real)kc.g[0]);
                vo = op.vi + dtr * (nc * ly + (Real)kc.g[1]);
                wo = op.wi + dtr * (nc * lz + (Real)kc.g[2]);
                xo = op.xi + dtr * uo; yo = op.yi + dtr * vo; zo = op.zi + dtr * wo;
            }
            b.x[i] = xo; b.y[i] = yo; b.z[i] = zo; b.u[i] = uo; b.v[i] = vo; b.w[i] = wo;
        }
    }
}

// ------------------------------------------------------------ collision ---
template <typename Real>
struct OpCol {
    const Real *x, *y, *z, *u, *v, *w;
    Real tc, rate;
    int i, sub, hits;
    bool active;
    Real xi, yi, zi, ui, vi, wi;
    Real cx, cy, cz;

    __device__ void load_cand(int j, Cand<Real>& c)
    {
        c.x = x[j]; c.y = y[j]; c.z = z[j]; c.j = j;
        c.a = u[j]; c.b = v[j]; c.c = w[j];
    }
    __device__ void process(const Cand<Real>* S, const uint16_t* L, int len)
    {
        for (int k = sub; k < len; k += TK) {
            const Cand<Real> c = S[L[k]];
            const Real dx = c.x - xi, dy = c.y - yi, dz = c.z - zi;
            const Real rr = rr3(dx, dy, dz);
            if (active && rr < tc && c.j != i) {
                // v = 0.5 (u_j - u_i) . r / |r|: its sign decides, formed as the reference forms it
                const Real hx = (Real)0.5 * (c.a - ui), hy = (Real)0.5 * (c.b - vi),
                           hz = (Real)0.5 * (c.c - wi);
                Real dot = hx * dx;
                dot = dot + hy * dy;
                dot = dot + hz * dz;
                if (dot <= (Real)0) {
                    const Real ods = rsq(rr);
                    const Real f = rate * (dot * ods) * ods;
                    cx = fmar(f, dx, cx);
                    cy = fmar(f, dy, cy);
                    cz = fmar(f, dz, cz);
                    hits++;
                }
            }
        }
    }
};

template <typename Real>
__global__ void __launch_bounds__(TB) k_collision(Grid g, const int* __restrict__ cs, StateView<Real> a,
                                                  StateView<Real> b, Thresh<Real> th, StepConst kc,
                                                  DevScalars* __restrict__ sc)
{
    __shared__ TileShared<Real> sh;
    const int cell = xcd_cell(g.ncells);
    if (cell >= g.ncells) return;
    const int s0 = cs[cell], s1 = cs[cell + 1];
    const double dt = sc->dt;
    const Lane ln;
    OpCol<Real> op;
    op.x = a.x; op.y = a.y; op.z = a.z; op.u = a.u; op.v = a.v; op.w = a.w;
    op.tc = th.t_col; op.rate = (Real)kc.col_rate; op.sub = ln.sub;
    op.hits = 0;
    const Real R2 = filt_r2(th.col);
    for (int c0 = s0; c0 < s1; c0 += TW * 8) {
        const int i = ln.particle(c0);
        const bool in = i < s1;
        op.i = i;
        op.active = false;
        op.cx = op.cy = op.cz = (Real)0;
        op.xi = op.yi = op.zi = op.ui = op.vi = op.wi = (Real)0;
        if (in) {
            op.xi = a.x[i]; op.yi = a.y[i]; op.zi = a.z[i];
            op.ui = a.u[i]; op.vi = a.v[i]; op.wi = a.w[i];
            op.active = a.type[i] == 0;
        }
        Real wlo[3], whi[3], blo[3], bhi[3];
        bool tact;
        if (tile_boxes(sh, op.active, op.xi, op.yi, op.zi, wlo, whi, blo, bhi, tact))
            tile_stage(g, cs, cell, sh, blo, bhi, wlo, whi, R2, tact, op);
        const Real cx = sub_sum(op.cx), cy = sub_sum(op.cy), cz = sub_sum(op.cz);
        if (in && ln.sub == 0) {
            const Real dtr = (Real)dt;
            b.u[i] = op.ui + cx; b.v[i] = op.vi + cy; b.w[i] = op.wi + cz;
            b.x[i] = op.xi + dtr * cx; b.y[i] = op.yi + dtr * cy; b.z[i] = op.zi + dtr * cz;
        }
    }
    const int hits = wave_sum(op.hits);
    if ((threadIdx.x & 63) == 0 && hits) atomicAdd(&sc->ncol, hits);
}

// ------------------------------------------------------ pressure gradient ---
template <typename Real>
struct OpGrad {
    const Real *x, *y, *z, *p;
    const int* type;
    Real re, tn, re2;
    int i, sub, pass, skip_dummy;
    bool active;
    Real xi, yi, zi, ph;
    Real gx, gy, gz;

    __device__ void load_cand(int j, Cand<Real>& c)
    {
        c.x = x[j]; c.y = y[j]; c.z = z[j]; c.j = j;
        c.t = type[j]; c.a = p[j];
    }
    __device__ void process(const Cand<Real>* S, const uint16_t* L, int len)
    {
        if (pass == 0) {
            // p_hat = min(p_i, p_j : type_j != 2, |r|^2 <= R0^2)
            for (int k = sub; k < len; k += TK) {
                const Cand<Real>& c = S[L[k]];
                const Real cx = c.x, cy = c.y, cz = c.z, pk = c.a;
                const int tk = c.t;
                const Real rr = rr3(cx - xi, cy - yi, cz - zi);
                const bool near = tk != 2 && rr <= re2 && pk < ph;
                ph = near ? pk : ph;
            }
            return;
        }
        for (int k = sub; k < len; k += TK) {
            const Cand<Real> c = S[L[k]];
            const Real dx = c.x - xi, dy = c.y - yi, dz = c.z - zi;
            const Real rr = rr3(dx, dy, dz);
            const bool inside = rr < tn && c.j != i && !(skip_dummy && c.t == 2);
            const Real ri = rsq(rr);
            const Real f = inside ? (c.a - ph) * (ri * ri) * fmar(re, ri, (Real)-1) : (Real)0;
            gx = fmar(f, dx, gx);
            gy = fmar(f, dy, gy);
            gz = fmar(f, dz, gz);
        }
    }
};

template <typename Real>
__global__ void __launch_bounds__(TB) k_gradp(Grid g, const int* __restrict__ cs, StateView<Real> a,
                                              StateView<Real> b, Thresh<Real> th, StepConst kc,
                                              DevScalars* __restrict__ sc)
{
    __shared__ TileShared<Real> sh;
    const int cell = xcd_cell(g.ncells);
    if (cell >= g.ncells) return;
    const int s0 = cs[cell], s1 = cs[cell + 1];
    const double dt = sc->dt, n0 = sc->n0;
    const Lane ln;
    OpGrad<Real> op;
    op.x = a.x; op.y = a.y; op.z = a.z; op.p = a.p; op.type = a.type;
    op.re = th.re_n; op.tn = th.t_n; op.re2 = th.re_n2; op.skip_dummy = kc.gradp_skip_dummy;
    op.sub = ln.sub;
    const Real R2 = filt_r2(th.re_n);
    float umax2 = 0.0f;
    for (int c0 = s0; c0 < s1; c0 += TW * 8) {
        const int i = ln.particle(c0);
        const bool in = i < s1;
        op.i = i;
        op.active = false;
        op.pass = 0;
        op.gx = op.gy = op.gz = (Real)0;
        op.xi = op.yi = op.zi = op.ph = (Real)0;
        Real ui = 0, vi = 0, wi = 0;
        if (in) {
            op.xi = a.x[i]; op.yi = a.y[i]; op.zi = a.z[i];
            ui = a.u[i]; vi = a.v[i]; wi = a.w[i];
            op.ph = a.p[i];
            op.active = a.type[i] == 0;
        }
        Real wlo[3], whi[3], blo[3], bhi[3];
        bool tact;
        if (tile_boxes(sh, op.active, op.xi, op.yi, op.zi, wlo, whi, blo, bhi, tact)) {
            const TileStage r = tile_stage(g, cs, cell, sh, blo, bhi, wlo, whi, R2, tact, op);
            op.ph = sub_min(op.ph);
            op.pass = 1;
            if (!r.overflowed) {
                if (tact) op.process(sh.S, sh.L[ln.wave], r.len);
                __syncthreads();
            } else {
                tile_stage(g, cs, cell, sh, blo, bhi, wlo, whi, R2, tact, op);
            }
        }
        const Real gx = sub_sum(op.gx), gy = sub_sum(op.gy), gz = sub_sum(op.gz);
        if (in && ln.sub == 0) {
            Real xo = op.xi, yo = op.yi, zo = op.zi, uo = ui, vo = vi, wo = wi;
            if (op.active) {
                const Real s = (Real)(-dt / kc.rho * (double)kc.dim / n0);
                const Real du = s * gx, dv = s * gy, dw = s * gz;
                const Real dtr = (Real)dt;
                uo = ui + du; vo = vi + dv; wo = wi + dw;
                xo = op.xi + dtr * du; yo = op.yi + dtr * dv; zo = op.zi + dtr * dw;
                const float m2 = (float)(uo * uo + vo * vo + wo * wo);
                umax2 = m2 > umax2 ? m2 : umax2;
            }
            b.x[i] = xo; b.y[i] = yo; b.z[i] = zo; b.u[i] = uo; b.v[i] = vo; b.w[i] = wo;
        }
    }
    umax2 = wave_max(umax2);
    if ((threadIdx.x & 63) == 0 && umax2 > 0.0f) atomicMax(&sc->maxu2_bits, __float_as_uint(umax2));
}

// ------------------------------------------------------------------- dt ---
// MPS_GPU::update_dt (MPS_GPU.cpp:761-777): dt = min(dt_max, CFL dp / max|u|);
// the flag replaces the reference's exit(1).  time accumulates dt.
__global__ void k_dt(StepConst kc, DevScalars* sc)
{
    if (threadIdx.x != 0) return;
    double m2 = (double)__uint_as_float(sc->maxu2_bits);
    if (kc.umax2_static > m2) m2 = kc.umax2_static;
    if (kc.mov_amp != 0.0) {
        // the moving walls carry A w cos(w t) from the last motion update
        const double va = (double)(float)(kc.mov_amp * kc.mov_omega * cos(kc.mov_omega * sc->time));
        if (va * va > m2) m2 = va * va;
    }
    double dt = kc.dt_max;
    if (m2 > 0.0) {
        const double t = kc.cfl * kc.dp / sqrt(m2);
        if (t < dt) dt = t;
    }
    if (dt < kc.dt_min) sc->err = 1;
    sc->dt = dt;
    sc->time += dt;
    sc->maxu2_bits = 0u;
    sc->ncol = 0;
}

// -------------------------------------------------------------- calOnce ---
// MPS::calOnce (MPS_CAL_expl_v1.0/MPS.cpp:311-451)
__global__ void __launch_bounds__(256) k_nmax(int np, const double* __restrict__ n, DevScalars* sc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) atomicMax(&sc->nmax_bits, (unsigned long long)__double_as_longlong(n[i]));
}
__global__ void __launch_bounds__(256) k_nargmax(int np, const double* __restrict__ n, DevScalars* sc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np && (unsigned long long)__double_as_longlong(n[i]) == sc->nmax_bits) atomicMin(&sc->argmax, i);
}

// lambda = sum_{j != i*} |r|^2 w(|r|; R0) / n0 around the first maximum, then
// the Laplacian coefficient 2 d / (n0 lambda).
template <typename Real>
__global__ void __launch_bounds__(64) k_lambda(Grid g, const int* __restrict__ cs, StateView<Real> s,
                                               Thresh<Real> th, int dim, DevScalars* sc)
{
    const int am = sc->argmax;
    const double n0 = __longlong_as_double((long long)sc->nmax_bits);
    int cell = 0;
    {   // cell of the argmax particle: last c with cs[c] <= am
        int lo = 0, hi = g.ncells - 1;
        while (lo < hi) { int mid = (lo + hi + 1) >> 1; if (cs[mid] <= am) lo = mid; else hi = mid - 1; }
        cell = lo;
    }
    int cx, cy, cz;
    cell_xyz(g, cell, cx, cy, cz);
    const Real xi = s.x[am], yi = s.y[am], zi = s.z[am];
    double acc = 0.0;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                const int xx = cx + dx, yy = cy + dy, zz = cz + dz;
                if (xx < 0 || yy < 0 || zz < 0 || xx >= g.dx || yy >= g.dy || zz >= g.dz) continue;
                const int c = zz * g.sheet + yy * g.dx + xx;
                for (int j = cs[c] + (int)threadIdx.x; j < cs[c + 1]; j += WAVE) {
                    if (j == am) continue;
                    const Real ex = s.x[j] - xi, ey = s.y[j] - yi, ez = s.z[j] - zi;
                    if (!(rr3(ex, ey, ez) < th.t_n)) continue;
                    const double fx = (double)s.x[j] - (double)xi, fy = (double)s.y[j] - (double)yi,
                                 fz = (double)s.z[j] - (double)zi;
                    const double r2 = fx * fx + fy * fy + fz * fz;
                    const double r = sqrt(r2);
                    acc += r2 * ((double)th.re_n / r - 1.0);
                }
            }
    acc = wave_sum(acc);
    if (threadIdx.x == 0) {
        sc->n0 = n0;
        sc->lambda = acc / n0;
        sc->coef = 2.0 * dim / (n0 * sc->lambda);
    }
}

// i_normal: nearest type-1 particle with |r|^2 < 2 R0^2, first strict minimum
// in index order (MPS_CAL_expl_v1.0/MPS.cpp:423-444); -1 when there is none.
template <typename Real>
__global__ void __launch_bounds__(256) k_normals(Grid g, const int* __restrict__ cs, int np,
                                                 StateView<Real> s, Thresh<Real> th)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    if (s.type[i] != 2) { s.normal[i] = -1; return; }
    int lo = 0, hi = g.ncells - 1;
    while (lo < hi) { int mid = (lo + hi + 1) >> 1; if (cs[mid] <= i) lo = mid; else hi = mid - 1; }
    int cx, cy, cz;
    cell_xyz(g, lo, cx, cy, cz);
    const Real xi = s.x[i], yi = s.y[i], zi = s.z[i];
    Real best = ((Real)2 * th.re_n) * th.re_n;
    int bj = -1;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                const int xx = cx + dx, yy = cy + dy, zz = cz + dz;
                if (xx < 0 || yy < 0 || zz < 0 || xx >= g.dx || yy >= g.dy || zz >= g.dz) continue;
                const int c = zz * g.sheet + yy * g.dx + xx;
                for (int j = cs[c]; j < cs[c + 1]; j++) {
                    if (s.type[j] != 1) continue;
                    const Real rr = rr3(s.x[j] - xi, s.y[j] - yi, s.z[j] - zi);
                    if (rr < best) { best = rr; bj = j; }
                }
            }
    s.normal[i] = bj;
}

// ------------------------------------------------------- neighbour lists ---
// Within-cutoff neighbour lists in the reference visiting order (ascending j),
// which = 0: R0, 1: RLAP.  fill = 0 counts into row_ptr[i + 1].
template <typename Real>
__global__ void __launch_bounds__(256) k_nlist(Grid g, const int* __restrict__ cs, int np,
                                               StateView<Real> s, Real t, int* __restrict__ row_ptr,
                                               int* __restrict__ col, int fill)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    int lo = 0, hi = g.ncells - 1;
    while (lo < hi) { int mid = (lo + hi + 1) >> 1; if (cs[mid] <= i) lo = mid; else hi = mid - 1; }
    int cx, cy, cz;
    cell_xyz(g, lo, cx, cy, cz);
    const Real xi = s.x[i], yi = s.y[i], zi = s.z[i];
    int k = fill ? row_ptr[i] : 0;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                const int xx = cx + dx, yy = cy + dy, zz = cz + dz;
                if (xx < 0 || yy < 0 || zz < 0 || xx >= g.dx || yy >= g.dy || zz >= g.dz) continue;
                const int c = zz * g.sheet + yy * g.dx + xx;
                for (int j = cs[c]; j < cs[c + 1]; j++) {
                    if (j == i) continue;
                    if (rr3(s.x[j] - xi, s.y[j] - yi, s.z[j] - zi) < t) {
                        if (fill) col[k] = j;
                        k++;
                    }
                }
            }
    if (!fill) row_ptr[i + 1] = k;
}

// ------------------------------------------------------------- launchers ---
static inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

template <typename Real>
void launch_pnd(const Grid& g, const int* cs, StateView<Real> s, const Thresh<Real>& th,
                const StepConst& k, DevScalars* sc, int all_types, int with_pressure, hipStream_t st)
{
    if (k.variant != 1) { wv::launch_pnd<Real>(g, cs, s, th, k, sc, all_types, with_pressure, st); return; }
    hipLaunchKernelGGL(k_pnd<Real>, dim3(xcd_grid(g.ncells)), dim3(TB), 0, st, g, cs, s, th, sc,
                       k.rho * k.cs * k.cs, all_types, with_pressure);
}

template <typename Real>
void launch_dummy_p(int n, StateView<Real> s, hipStream_t st)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_dummy_p<Real>, dim3(cdiv(n, 256)), dim3(256), 0, st, n, s);
}

template <typename Real>
void launch_visc(const Grid& g, const int* cs, StateView<Real> a, StateView<Real> b,
                 const Thresh<Real>& th, const StepConst& k, DevScalars* sc, hipStream_t st)
{
    if (k.variant != 1) { wv::launch_visc<Real>(g, cs, a, b, th, k, sc, st); return; }
    hipLaunchKernelGGL(k_visc<Real>, dim3(xcd_grid(g.ncells)), dim3(TB), 0, st, g, cs, a, b, th, k, sc);
}

template <typename Real>
void launch_collision(const Grid& g, const int* cs, StateView<Real> a, StateView<Real> b,
                      const Thresh<Real>& th, const StepConst& k, DevScalars* sc, hipStream_t st)
{
    if (k.variant != 1) { wv::launch_collision<Real>(g, cs, a, b, th, k, sc, st); return; }
    hipLaunchKernelGGL(k_collision<Real>, dim3(xcd_grid(g.ncells)), dim3(TB), 0, st, g, cs, a, b, th, k, sc);
}

template <typename Real>
void launch_gradp(const Grid& g, const int* cs, StateView<Real> a, StateView<Real> b,
                  const Thresh<Real>& th, const StepConst& k, DevScalars* sc, hipStream_t st)
{
    if (k.variant != 1) { wv::launch_gradp<Real>(g, cs, a, b, th, k, sc, st); return; }
    hipLaunchKernelGGL(k_gradp<Real>, dim3(xcd_grid(g.ncells)), dim3(TB), 0, st, g, cs, a, b, th, k, sc);
}

void launch_dt(const StepConst& k, DevScalars* sc, hipStream_t st)
{
    hipLaunchKernelGGL(k_dt, dim3(1), dim3(64), 0, st, k, sc);
}

// n over every particle, n0 = first maximum, lambda and coef, normals.
// The caller zeroes nmax_bits and sets argmax to INT_MAX beforehand.
template <typename Real>
void launch_calonce(const Grid& g, const int* cs, int np, StateView<Real> s, const Thresh<Real>& th,
                    const StepConst& k, DevScalars* sc, hipStream_t st)
{
    launch_pnd<Real>(g, cs, s, th, k, sc, 1, 0, st);
    if (np > 0) {
        hipLaunchKernelGGL(k_nmax, dim3(cdiv(np, 256)), dim3(256), 0, st, np, s.n, sc);
        hipLaunchKernelGGL(k_nargmax, dim3(cdiv(np, 256)), dim3(256), 0, st, np, s.n, sc);
        hipLaunchKernelGGL(k_lambda<Real>, dim3(1), dim3(64), 0, st, g, cs, s, th, k.dim, sc);
        hipLaunchKernelGGL(k_normals<Real>, dim3(cdiv(np, 256)), dim3(256), 0, st, g, cs, np, s, th);
    }
}

template <typename Real>
void launch_nlist(const Grid& g, const int* cs, int np, StateView<Real> s, const Thresh<Real>& th,
                  int which, int* row_ptr, int* col, int fill, hipStream_t st)
{
    if (np == 0) return;
    hipLaunchKernelGGL(k_nlist<Real>, dim3(cdiv(np, 256)), dim3(256), 0, st, g, cs, np, s,
                       which ? th.t_lap : th.t_n, row_ptr, col, fill);
}

#define MPS_INST(R)                                                                                   \
    template void launch_pnd<R>(const Grid&, const int*, StateView<R>, const Thresh<R>&,            \
                                const StepConst&, DevScalars*, int, int, hipStream_t);               \
    template void launch_dummy_p<R>(int, StateView<R>, hipStream_t);                                 \
    template void launch_visc<R>(const Grid&, const int*, StateView<R>, StateView<R>,               \
                                 const Thresh<R>&, const StepConst&, DevScalars*, hipStream_t);      \
    template void launch_collision<R>(const Grid&, const int*, StateView<R>, StateView<R>,          \
                                      const Thresh<R>&, const StepConst&, DevScalars*, hipStream_t); \
    template void launch_gradp<R>(const Grid&, const int*, StateView<R>, StateView<R>,              \
                                  const Thresh<R>&, const StepConst&, DevScalars*, hipStream_t);     \
    template void launch_calonce<R>(const Grid&, const int*, int, StateView<R>, const Thresh<R>&,   \
                                    const StepConst&, DevScalars*, hipStream_t);                     \
    template void launch_nlist<R>(const Grid&, const int*, int, StateView<R>, const Thresh<R>&, int, \
                                  int*, int*, int, hipStream_t);
MPS_INST(float)
MPS_INST(double)
#undef MPS_INST

}  // namespace mps
