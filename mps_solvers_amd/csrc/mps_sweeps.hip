// mps_sweeps.hip -- the small kernels of the MPS step on gfx950: the dummy
// pressure copy, the time step, calOnce (n0, lambda, normals) and the
// within-cutoff neighbour lists used by the parity tests, plus the launchers
// of the neighbour sweeps, which live in mps_sweeps_wave.hip.
//
// Operators of the sweeps (reference semantics, two-pass as in MPS_CAL_expl_v1.0):
//   PND + explicit pressure  MPS_GPU::cal_n (MPS_GPU.cpp:830-887) + calPres_expl
//                            (MPS_GPU_EXPL.cpp:199-238): n in FP64, p = max(0, rho c^2 (n-n0)/n0)
//   viscosity + gravity      calVisc_expl (MPS_GPU_EXPL.cpp:152-194), d3_LapVel (MPS_GPU.cpp:683-717)
//   collision                MPS_GPU::collision (MPS_GPU.cpp:892-986)
//   pressure gradient        MPS_GPU::calDash (MPS_GPU.cpp:782-825), d3_GradPres (MPS_GPU.cpp:571-638)
#include <float.h>

#include "mps_device.hpp"
#include "mps_kernels.hpp"

namespace mps {

// dummy pressure p_i = p_normal(i) (MPS_GPU_EXPL.cpp:228-236); no partner -> 0
template <typename Real>
__global__ void __launch_bounds__(256) k_dummy_p(int n, StateView<Real> s)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || s.type[i] != 2) return;
    const int m = s.normal[i];
    s.p[i] = m >= 0 ? s.p[m] : (Real)0;
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
    wv::launch_pnd<Real>(g, cs, s, th, k, sc, all_types, with_pressure, st);
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
    wv::launch_visc<Real>(g, cs, a, b, th, k, sc, st);
}

template <typename Real>
void launch_collision(const Grid& g, const int* cs, StateView<Real> a, StateView<Real> b,
                      const Thresh<Real>& th, const StepConst& k, DevScalars* sc, hipStream_t st)
{
    wv::launch_collision<Real>(g, cs, a, b, th, k, sc, st);
}

template <typename Real>
void launch_gradp(const Grid& g, const int* cs, StateView<Real> a, StateView<Real> b,
                  const Thresh<Real>& th, const StepConst& k, DevScalars* sc, hipStream_t st)
{
    wv::launch_gradp<Real>(g, cs, a, b, th, k, sc, st);
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
