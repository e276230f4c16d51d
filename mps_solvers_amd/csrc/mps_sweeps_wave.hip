// mps_sweeps_wave.hip -- the one-wavefront-per-cell neighbour sweeps: one
// 64-lane wavefront per cell (chunks of 64 particles, lane = particle), the
// 27-cell candidate set filtered against the wave's bounding box and staged
// through LDS in batches, each staged candidate broadcast to all lanes.  The
// viscosity sweep splits each chunk into two lane groups (two lanes per
// particle) with a candidate set per group.
#include <float.h>

#include "mps_device.hpp"
#include "mps_kernels.hpp"
#include "mps_stage_wave.hpp"

namespace mps {
namespace wv {

// ------------------------------------------------------------------ PND ---
// Two phases per LDS batch: the FP32 cutoff test appends the in-range
// candidates to a per-lane list (slot-major in LDS, so that every lane's k-th
// entry sits in its own bank), then the FP64 weight runs over the dense
// lists only.  Each lane still adds its neighbours in ascending staged order.
constexpr int PND_LMAX = 32;   // list slots per lane (drained before any lane can overflow)

template <typename Real>
struct OpPnd {
    const Real *x, *y, *z;
    const int* type;
    Real re;     // R0 in the state precision
    Real tn;     // rr < tn  <=>  sqrt(rr) < R0
    int all_types;
    int i;
    bool active;
    Real xi, yi, zi;
    double acc;
    uint16_t* L;   // L[s * WAVE + lane]

    __device__ void load_self(int i_, bool in)
    {
        i = i_;
        acc = 0.0;
        active = false;
        if (in) {
            xi = x[i]; yi = y[i]; zi = z[i];
            active = all_types || type[i] != 2;
        }
    }
    __device__ void load_cand(int j, CandP<Real>& c) { c.x = x[j]; c.y = y[j]; c.z = z[j]; c.j = j; }
    __device__ void drain(const CandP<Real>* buf, int m)
    {
        const int lane = threadIdx.x;
        const int mm = wave_max(m);
        const double red = (double)re, xd = (double)xi, yd = (double)yi, zd = (double)zi;
        for (int s = 0; s < mm; s++) {
            if (s < m) {
                const CandP<Real> c = buf[L[s * WAVE + lane]];
                // exact differences of the stored values (FP64 holds any float difference)
                const double ddx = (double)c.x - xd, ddy = (double)c.y - yd, ddz = (double)c.z - zd;
                const double r2 = __builtin_fma(ddx, ddx, __builtin_fma(ddy, ddy, ddz * ddz));
                acc += __builtin_fma(red, rsq(r2), -1.0);
            }
        }
    }
    __device__ __forceinline__ void test(const CandP<Real> c, int k, int& m)
    {
        // branch-free append: the slot is written either way (m <= PND_LMAX - 1
        // here, see process) and only a kept candidate advances m
        const Real rr = rr3(c.x - xi, c.y - yi, c.z - zi);
        const bool take = active & (rr < tn) & (c.j != i);
        L[m * WAVE + threadIdx.x] = (uint16_t)k;
        m += take ? 1 : 0;
    }
    __device__ void process(const CandP<Real>* buf, int cnt)
    {
        int m = 0;
        for (int k0 = 0; k0 < cnt; k0 += 16) {
            for_each_cand<8>(buf + k0, min(16, cnt - k0), [&](const CandP<Real>& c, int k) { test(c, k0 + k, m); });
            // at most 16 appends between checks: no lane passes PND_LMAX
            if (ballot(m > PND_LMAX - 16)) {
                drain(buf, m);
                m = 0;
            }
        }
        drain(buf, m);
    }
};

template <typename Real, int CAP>
__global__ void __launch_bounds__(64) k_pnd(Grid g, const int* __restrict__ cs, StateView<Real> s,
                                            Thresh<Real> th, const DevScalars* __restrict__ sc,
                                            double one_over_alpha, int all_types, int with_pressure)
{
    __shared__ CandP<Real> buf[CAP];
    __shared__ uint16_t L[PND_LMAX * WAVE];
    const int cell = xcd_cell(g.ncells);
    if (cell >= g.ncells) return;
    const int s0 = cs[cell], s1 = cs[cell + 1];
    const double n0 = with_pressure ? sc->n0 : 1.0;
    OpPnd<Real> op;
    op.x = s.x; op.y = s.y; op.z = s.z; op.type = s.type;
    op.re = th.re_n; op.tn = th.t_n; op.all_types = all_types;
    op.L = L;
    const Real R2 = filt_r2(th.re_n);
    for (int c0 = s0; c0 < s1; c0 += WAVE) {
        const int i = c0 + (int)threadIdx.x;
        const bool in = i < s1;
        op.load_self(i, in);
        Real lo[3], hi[3];
        wave_box(op.active, op.xi, op.yi, op.zi, lo, hi);
        if (!ballot(op.active)) continue;
        stage<CAP>(g, cs, cell, lo, hi, R2, buf, op);
        if (op.active) {
            s.n[i] = op.acc;
            if (with_pressure) {
                const double t = one_over_alpha * (op.acc - n0) / n0;
                s.p[i] = (Real)(t > 0.0 ? t : 0.0);
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ viscosity ---
// sum_j w_ij (u_j - u_i) = sum_j w_ij (u_j - ub) - (u_i - ub) sum_j w_ij with ub
// the chunk's mean velocity: u_j - ub is formed once per staged candidate, so
// a visit costs three FMAs and one add instead of three differences and
// three FMAs.  The differences stay relative to the local flow (no added
// cancellation against the absolute velocity).
//
// The chunk is cut into VISC_GROUPS lane groups (two halves at the median of
// the longest axis), each staging its own, smaller, candidate set and walking
// it with two lanes per particle, two candidates read ahead.  At the 4 l0
// radius this evaluates ~23% fewer candidates per particle than one set for
// the whole chunk; four groups pay more in staging than they save.
constexpr int VISC_GROUPS = 2, VISC_READ_AHEAD = 2;

template <typename Real, int G>
struct OpVisc {
    const Real *x, *y, *z, *u, *v, *w;
    Real re, tl;
    int i, q;
    Real xi, yi, zi;
    Real ub, vb, wb;      // chunk-mean velocity
    Real lx, ly, lz, sw;

    __device__ void load_cand(int j, Cand<Real>& c)
    {
        c.x = x[j]; c.y = y[j]; c.z = z[j]; c.j = j;
        c.a = u[j] - ub; c.b = v[j] - vb; c.c = w[j] - wb;
    }
    __device__ __forceinline__ void visit(const Cand<Real> c)
    {
        const Real dx = c.x - xi, dy = c.y - yi, dz = c.z - zi;
        const Real rr = rr3(dx, dy, dz);
        const bool inside = (rr < tl) & (c.j != i);
        const Real wt = inside ? fmar(re, rsq(rr), (Real)-1) : (Real)0;
        sw += wt;
        lx = fmar(wt, c.a, lx);
        ly = fmar(wt, c.b, ly);
        lz = fmar(wt, c.c, lz);
    }
    __device__ void process(Cand<Real>* buf, int cnt)
    {
        const int n = pad_phases<G>(buf, cnt);
        for_each_phase<VISC_READ_AHEAD, G>(buf, 0, n, q, [&](const Cand<Real>& c, int) { visit(c); });
    }
};

template <typename Real, int CAP, int G>
__global__ void __launch_bounds__(64) k_visc(Grid g, const int* __restrict__ cs, StateView<Real> a,
                                             StateView<Real> b, Thresh<Real> th, StepConst kc,
                                             const DevScalars* __restrict__ sc)
{
    constexpr int S = WAVE / G;
    __shared__ Cand<Real> buf[CAP + G];
    __shared__ int ord[WAVE];
    __shared__ Real sums[4][WAVE];   // per chunk particle: sum w, sum w (u_j - ub)
    const int cell = xcd_cell(g.ncells);
    if (cell >= g.ncells) return;
    const int s0 = cs[cell], s1 = cs[cell + 1];
    const double dt = sc->dt, coef = sc->coef;
    const int lane = threadIdx.x, sl = lane % S;
    OpVisc<Real, G> op;
    op.x = a.x; op.y = a.y; op.z = a.z; op.u = a.u; op.v = a.v; op.w = a.w;
    op.re = th.re_lap; op.tl = th.t_lap;
    op.q = lane / S;
    const Real R2 = filt_r2(th.re_lap);
    for (int c0 = s0; c0 < s1; c0 += WAVE) {
        const int i = c0 + lane;
        bool act = false;
        Real xi = 0, yi = 0, zi = 0, ui = 0, vi = 0, wi = 0;
        if (i < s1) {
            xi = a.x[i]; yi = a.y[i]; zi = a.z[i];
            ui = a.u[i]; vi = a.v[i]; wi = a.w[i];
            act = a.type[i] == 0;
            if (!act) {   // walls and dummies keep their state
                b.x[i] = xi; b.y[i] = yi; b.z[i] = zi; b.u[i] = ui; b.v[i] = vi; b.w[i] = wi;
            }
        }
        const uint64_t am = ballot(act);
        if (!am) continue;
        Real lo[3], hi[3];
        wave_box(act, xi, yi, zi, lo, hi);
        const Real inv = (Real)1 / (Real)__popcll(am);
        op.ub = wave_sum(act ? ui : (Real)0) * inv;
        op.vb = wave_sum(act ? vi : (Real)0) * inv;
        op.wb = wave_sum(act ? wi : (Real)0) * inv;
        const int ng = split_groups<G>(act, xi, yi, zi, lo, hi, ord);
        for (int gi = 0; gi < ng; gi++) {
            const int o = ord[gi * S + sl];
            const bool has = o >= 0;
            const int j = c0 + (has ? o : 0);
            op.i = j;
            op.xi = a.x[j]; op.yi = a.y[j]; op.zi = a.z[j];
            op.lx = op.ly = op.lz = op.sw = (Real)0;
            Real glo[3], ghi[3];
            wave_box(has, op.xi, op.yi, op.zi, glo, ghi);
            stage<CAP>(g, cs, cell, glo, ghi, R2, buf, op);
            const Real sw = phase_sum<G>(op.sw);
            const Real lx = phase_sum<G>(op.lx), ly = phase_sum<G>(op.ly), lz = phase_sum<G>(op.lz);
            if (has && op.q == 0) {
                sums[0][o] = sw; sums[1][o] = lx; sums[2][o] = ly; sums[3][o] = lz;
            }
            __syncthreads();
        }
        // each particle's own lane finishes it: one coalesced store per array
        // (its state is read again rather than held in registers across the groups)
        if (act) {
            const Real sw = sums[0][lane], lx = sums[1][lane], ly = sums[2][lane], lz = sums[3][lane];
            xi = a.x[i]; yi = a.y[i]; zi = a.z[i];
            ui = a.u[i]; vi = a.v[i]; wi = a.w[i];
            const Real dtr = (Real)dt, nc = (Real)(kc.nu * coef);
            const Real tx = fmar(-(ui - op.ub), sw, lx);
            const Real ty = fmar(-(vi - op.vb), sw, ly);
            const Real tz = fmar(-(wi - op.wb), sw, lz);
            const Real uo = ui + dtr * (nc * tx + (Real)kc.g[0]);
            const Real vo = vi + dtr * (nc * ty + (Real)kc.g[1]);
            const Real wo = wi + dtr * (nc * tz + (Real)kc.g[2]);
            b.x[i] = xi + dtr * uo; b.y[i] = yi + dtr * vo; b.z[i] = zi + dtr * wo;
            b.u[i] = uo; b.v[i] = vo; b.w[i] = wo;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ collision ---
template <typename Real>
struct OpCol {
    const Real *x, *y, *z, *u, *v, *w;
    Real tc, rate;
    int i, hits;
    bool active;
    Real xi, yi, zi, ui, vi, wi;
    Real cx, cy, cz;

    __device__ void load_cand(int j, Cand<Real>& c)
    {
        c.x = x[j]; c.y = y[j]; c.z = z[j]; c.j = j;
        c.a = u[j]; c.b = v[j]; c.c = w[j];
    }
    __device__ __forceinline__ void visit(const Cand<Real> c)
    {
        const Real dx = c.x - xi, dy = c.y - yi, dz = c.z - zi;
        const Real rr = rr3(dx, dy, dz);
        if (rr < tc && c.j != i) {
            // v = 0.5 (u_j - u_i) . r / |r|; the sign decides, formed as the reference forms it
            const Real hx = (Real)0.5 * (c.a - ui), hy = (Real)0.5 * (c.b - vi), hz = (Real)0.5 * (c.c - wi);
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
    __device__ void process(const Cand<Real>* buf, int cnt)
    {
        for_each_cand<1>(buf, cnt, [&](const Cand<Real>& c, int) { visit(c); });
    }
};

template <typename Real, int CAP>
__global__ void __launch_bounds__(64) k_collision(Grid g, const int* __restrict__ cs, StateView<Real> a,
                                                  StateView<Real> b, Thresh<Real> th, StepConst kc,
                                                  DevScalars* __restrict__ sc)
{
    __shared__ Cand<Real> buf[CAP];
    const int cell = xcd_cell(g.ncells);
    if (cell >= g.ncells) return;
    const int s0 = cs[cell], s1 = cs[cell + 1];
    const double dt = sc->dt;
    OpCol<Real> op;
    op.x = a.x; op.y = a.y; op.z = a.z; op.u = a.u; op.v = a.v; op.w = a.w;
    op.tc = th.t_col; op.rate = (Real)kc.col_rate;
    const Real R2 = filt_r2(th.col);
    int hits = 0;
    for (int c0 = s0; c0 < s1; c0 += WAVE) {
        const int i = c0 + (int)threadIdx.x;
        const bool in = i < s1;
        op.i = i;
        op.active = false;
        op.hits = 0;
        op.cx = op.cy = op.cz = (Real)0;
        if (in) {
            op.xi = a.x[i]; op.yi = a.y[i]; op.zi = a.z[i];
            op.ui = a.u[i]; op.vi = a.v[i]; op.wi = a.w[i];
            op.active = a.type[i] == 0;
        }
        Real lo[3], hi[3];
        wave_box(op.active, op.xi, op.yi, op.zi, lo, hi);
        if (ballot(op.active)) {
            stage<CAP, true>(g, cs, cell, lo, hi, R2, buf, op);
        }
        if (in) {
            Real cx = op.active ? op.cx : (Real)0, cy = op.active ? op.cy : (Real)0,
                 cz = op.active ? op.cz : (Real)0;
            const Real dtr = (Real)dt;
            b.u[i] = op.ui + cx; b.v[i] = op.vi + cy; b.w[i] = op.wi + cz;
            b.x[i] = op.xi + dtr * cx; b.y[i] = op.yi + dtr * cy; b.z[i] = op.zi + dtr * cz;
            if (op.active) hits += op.hits;
        }
        __syncthreads();
    }
    hits = wave_sum(hits);
    if (threadIdx.x == 0 && hits) atomicAdd(&sc->ncol, hits);
}

// ------------------------------------------------------ pressure gradient ---
// One pass over the candidates.  The reference's sum (MPS_GPU.cpp:571-638)
//   sum_j (p_j - p_hat) w_ij / r_ij^2 r_ij,  p_hat = min(p_i, p_j : type_j != 2, r^2 <= R0^2)
// needs p_hat before it starts; it equals
//   sum_j (p_j - p_i) w_ij / r_ij^2 r_ij  +  (p_i - p_hat) sum_j w_ij / r_ij^2 r_ij,
// so both sums and p_hat accumulate together and combine at the end.  The
// terms are the natural differences p_j - p_i; no cancellation is added.
template <typename Real>
struct OpGrad {
    const Real *x, *y, *z, *p;
    const int* type;
    Real re, tn, re2;
    int i, skip_dummy;
    bool active;
    Real xi, yi, zi, pi, ph;
    Real ax, ay, az, bx, by, bz;

    // c.a = p_j; c.b = p_j as a p-hat candidate (the largest value for a dummy,
    // which never lowers p-hat).  With skip_dummy a dummy adds nothing at all,
    // so it is placed out of every box and never staged.
    __device__ void load_cand(int j, Cand<Real>& c)
    {
        const bool dummy = type[j] == 2;
        c.x = (skip_dummy && dummy) ? big<Real>() : x[j];
        c.y = y[j]; c.z = z[j]; c.j = j;
        c.a = p[j];
        c.b = dummy ? big<Real>() : c.a;
    }
    __device__ __forceinline__ void visit(const Cand<Real> c)
    {
        const Real dx = c.x - xi, dy = c.y - yi, dz = c.z - zi;
        const Real rr = rr3(dx, dy, dz);
        ph = (rr <= re2) & (c.b < ph) ? c.b : ph;
        const bool inside = (rr < tn) & (c.j != i);
        const Real ri = rsq(rr);
        const Real s = inside ? (ri * ri) * fmar(re, ri, (Real)-1) : (Real)0;
        const Real f = s * (c.a - pi);
        ax = fmar(f, dx, ax);
        ay = fmar(f, dy, ay);
        az = fmar(f, dz, az);
        bx = fmar(s, dx, bx);
        by = fmar(s, dy, by);
        bz = fmar(s, dz, bz);
    }
    __device__ void process(const Cand<Real>* buf, int cnt)
    {
        for_each_cand<4>(buf, cnt, [&](const Cand<Real>& c, int) { visit(c); });
    }
};

template <typename Real, int CAP>
__global__ void __launch_bounds__(64) k_gradp(Grid g, const int* __restrict__ cs, StateView<Real> a,
                                              StateView<Real> b, Thresh<Real> th, StepConst kc,
                                              DevScalars* __restrict__ sc)
{
    __shared__ Cand<Real> buf[CAP];
    const int cell = xcd_cell(g.ncells);
    if (cell >= g.ncells) return;
    const int s0 = cs[cell], s1 = cs[cell + 1];
    const double dt = sc->dt, n0 = sc->n0;
    OpGrad<Real> op;
    op.x = a.x; op.y = a.y; op.z = a.z; op.p = a.p; op.type = a.type;
    op.re = th.re_n; op.tn = th.t_n; op.re2 = th.re_n2; op.skip_dummy = kc.gradp_skip_dummy;
    const Real R2 = filt_r2(th.re_n);
    float umax2 = 0.0f;
    for (int c0 = s0; c0 < s1; c0 += WAVE) {
        const int i = c0 + (int)threadIdx.x;
        const bool in = i < s1;
        op.i = i;
        op.active = false;
        op.ax = op.ay = op.az = op.bx = op.by = op.bz = (Real)0;
        Real ui = 0, vi = 0, wi = 0;
        if (in) {
            op.xi = a.x[i]; op.yi = a.y[i]; op.zi = a.z[i];
            ui = a.u[i]; vi = a.v[i]; wi = a.w[i];
            op.pi = a.p[i];
            op.ph = op.pi;
            op.active = a.type[i] == 0;
        }
        Real lo[3], hi[3];
        wave_box(op.active, op.xi, op.yi, op.zi, lo, hi);
        if (ballot(op.active)) stage<CAP>(g, cs, cell, lo, hi, R2, buf, op);
        if (in) {
            Real xo = op.xi, yo = op.yi, zo = op.zi, uo = ui, vo = vi, wo = wi;
            if (op.active) {
                const Real dp = op.pi - op.ph;
                const Real gx = fmar(dp, op.bx, op.ax), gy = fmar(dp, op.by, op.ay), gz = fmar(dp, op.bz, op.az);
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
        __syncthreads();
    }
    umax2 = wave_max(umax2);
    if (threadIdx.x == 0 && umax2 > 0.0f) atomicMax(&sc->maxu2_bits, __float_as_uint(umax2));
}

// LDS batches of 128 FP32 candidates (4 KiB; 64 FP64 candidates): VGPRs, not
// LDS, then bound the resident wavefronts (8 per SIMD).  256- and 512-candidate
// batches measured slower (fewer resident wavefronts).
constexpr int WAVE_CAP = 128;
#define MPS_WAVE_LAUNCH(kern, ...)                                                                          \
    hipLaunchKernelGGL((kern<Real, cap_of<Real>(WAVE_CAP)>), dim3(xcd_grid(g.ncells)), dim3(64), 0, st, \
                       __VA_ARGS__)
#define MPS_WAVE_LAUNCH_G(kern, G, ...)                                                                       \
    hipLaunchKernelGGL((kern<Real, cap_of<Real>(WAVE_CAP), G>), dim3(xcd_grid(g.ncells)), dim3(64), 0, st, \
                       __VA_ARGS__)

template <typename Real>
void launch_pnd(const Grid& g, const int* cs, StateView<Real> s, const Thresh<Real>& th, const StepConst& k,
                DevScalars* sc, int all_types, int with_pressure, hipStream_t st)
{
    MPS_WAVE_LAUNCH(k_pnd, g, cs, s, th, sc, k.rho * k.cs * k.cs, all_types, with_pressure);
}
template <typename Real>
void launch_visc(const Grid& g, const int* cs, StateView<Real> a, StateView<Real> b, const Thresh<Real>& th,
                 const StepConst& k, DevScalars* sc, hipStream_t st)
{
    MPS_WAVE_LAUNCH_G(k_visc, VISC_GROUPS, g, cs, a, b, th, k, sc);
}
template <typename Real>
void launch_collision(const Grid& g, const int* cs, StateView<Real> a, StateView<Real> b, const Thresh<Real>& th,
                      const StepConst& k, DevScalars* sc, hipStream_t st)
{
    MPS_WAVE_LAUNCH(k_collision, g, cs, a, b, th, k, sc);
}
template <typename Real>
void launch_gradp(const Grid& g, const int* cs, StateView<Real> a, StateView<Real> b, const Thresh<Real>& th,
                  const StepConst& k, DevScalars* sc, hipStream_t st)
{
    MPS_WAVE_LAUNCH(k_gradp, g, cs, a, b, th, k, sc);
}
#undef MPS_WAVE_LAUNCH
#undef MPS_WAVE_LAUNCH_G

#define MPS_INST_W(R)                                                                                          \
    template void launch_pnd<R>(const Grid&, const int*, StateView<R>, const Thresh<R>&, const StepConst&,   \
                                DevScalars*, int, int, hipStream_t);                                          \
    template void launch_visc<R>(const Grid&, const int*, StateView<R>, StateView<R>, const Thresh<R>&,      \
                                 const StepConst&, DevScalars*, hipStream_t);                                 \
    template void launch_collision<R>(const Grid&, const int*, StateView<R>, StateView<R>, const Thresh<R>&, \
                                      const StepConst&, DevScalars*, hipStream_t);                            \
    template void launch_gradp<R>(const Grid&, const int*, StateView<R>, StateView<R>, const Thresh<R>&,     \
                                  const StepConst&, DevScalars*, hipStream_t);
MPS_INST_W(float)
MPS_INST_W(double)
#undef MPS_INST_W

}  // namespace wv
}  // namespace mps
