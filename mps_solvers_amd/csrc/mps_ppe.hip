// mps_ppe.hip -- pressure Poisson equation of the semi-implicit MPS step:
// CSR assembly and a hand-written conjugate-gradient solver on gfx950.
//
// Assembly restates MPS_CPU::buildPoisson (MPS_CAL_v4.0/MPS_CPU.cpp:619-710):
//   bc_i = n_i < beta n0; rows with bc_i or type 2 are identity with b = 0;
//   otherwise A_ij = w(|r_ij|; RLAP) for j != i, type_j != 2, j not bc, and
//   A_ii = -sum_{j != i, type_j != 2} w_ij - rho lambda alpha n0 / (2 d dt^2),
//   b_i = lambda rho / (2 d dt^2) (n0 - n_i).
// Columns ascend (the reference's visiting order).  Values are formed in FP64
// and stored twice: FP32 for the inner CG SpMV, FP64 for the outer residual.
//
// The solver replaces MPS_CPU::DoCG (MPS_CPU.cpp:829-888) and its cuBLAS twin
// (MPS_CAL_v4.0/cuda_ker.cu:171-297): CG on the CSR matrix with FP32 SpMV
// inside an FP64 iterative-refinement loop (or plain FP64 CG).  Reductions are
// deterministic: per-block FP64 partials summed in a fixed order (in the CG
// loop by every block of the consuming kernel, in the once-per-system row
// setup by the last block to finish).  No kernel waits on another block.
//
// Assembly runs as two one-wavefront-per-cell sweeps over the staged 27-cell
// candidates (mps_stage_wave.hpp): a count pass for the row lengths, a scan,
// and a fill pass that writes each row's entries in ascending column order.
#include <algorithm>

#include "mps_device.hpp"
#include "mps_kernels.hpp"
#include "mps_ppe.hpp"
#include "mps_stage_wave.hpp"

namespace mps {

// ------------------------------------------------------------ row setup ---
template <typename Real>
__global__ void __launch_bounds__(256) k_ppe_rows(int n, StateView<Real> s, const DevScalars* __restrict__ sc,
                                                  double beta, double btmp_over, int x0_kind,
                                                  const double* __restrict__ x0_ext,
                                                  unsigned char* __restrict__ bc, double* __restrict__ b,
                                                  double* __restrict__ x, double* __restrict__ partial,
                                                  PpeScalars* ps)
{
    const double n0 = sc->n0, lambda = sc->lambda, dt = sc->dt;
    double bb = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double ni = s.n[i];
        const bool isbc = ni < beta * n0;
        const bool id = isbc || s.type[i] == 2;
        // lambda rho / (2 d dt^2) = lambda * btmp_over / dt^2
        const double bi = id ? 0.0 : lambda * btmp_over / (dt * dt) * (n0 - ni);
        bc[i] = isbc ? 1 : 0;
        b[i] = bi;
        double xi = 0.0;
        if (x0_kind == 1) xi = (double)s.p[i];
        else if (x0_kind == 2) xi = x0_ext[i];
        x[i] = isbc ? 0.0 : xi;
        bb += bi * bi;
    }
    block_sum_last(bb, partial, &ps->ticket[0], &ps->bb);
}

// ------------------------------------------------------------- assembly ---
// Row lengths for the CSR scan: one wavefront per cell, lane = row, each
// staged candidate broadcast to all lanes; only the FP32 cutoff test (the
// weights come in the fill pass).  A row counts its diagonal plus every j != i
// with sqrt(rr) < RLAP, type_j != 2 and j not Dirichlet; identity rows count 1.
constexpr int PPE_T2 = 1 << 30;   // candidate flag: type 2
constexpr int PPE_BC = 1 << 29;   // candidate flag: Dirichlet
constexpr int PPE_JM = PPE_BC - 1;

template <typename Real>
struct OpCount {
    const Real *x, *y, *z;
    const int* type;
    const unsigned char* bcv;
    Real tl;
    int i, count;
    bool active;
    Real xi, yi, zi;

    __device__ void load_cand(int j, wv::CandP<Real>& c)
    {
        c.x = x[j]; c.y = y[j]; c.z = z[j];
        c.j = j | (type[j] == 2 ? PPE_T2 : 0) | (bcv[j] ? PPE_BC : 0);
    }
    __device__ void process(const wv::CandP<Real>* buf, int cnt)
    {
        wv::for_each_cand<8>(buf, cnt, [&](const wv::CandP<Real>& c, int) {
            const Real rr = rr3(c.x - xi, c.y - yi, c.z - zi);
            count += ((rr < tl) & ((c.j & (PPE_T2 | PPE_BC)) == 0) & ((c.j & PPE_JM) != i)) ? 1 : 0;
        });
    }
};

template <typename Real, int CAP>
__global__ void __launch_bounds__(64) k_ppe_count(Grid g, const int* __restrict__ cs, StateView<Real> s,
                                                  Thresh<Real> th, const unsigned char* __restrict__ bc,
                                                  int* __restrict__ row_ptr)
{
    __shared__ wv::CandP<Real> buf[CAP];
    const int cell = xcd_cell(g.ncells);
    if (cell >= g.ncells) return;
    const int s0 = cs[cell], s1 = cs[cell + 1];
    OpCount<Real> op;
    op.x = s.x; op.y = s.y; op.z = s.z; op.type = s.type; op.bcv = bc;
    op.tl = th.t_lap;
    const Real R2 = wv::filt_r2(th.re_lap);
    for (int c0 = s0; c0 < s1; c0 += WAVE) {
        const int i = c0 + (int)threadIdx.x;
        const bool in = i < s1;
        op.i = i;
        op.count = 0;
        op.active = false;
        if (in) {
            op.xi = s.x[i]; op.yi = s.y[i]; op.zi = s.z[i];
            op.active = !bc[i] && s.type[i] != 2;
        }
        Real lo[3], hi[3];
        wv::wave_box(op.active, op.xi, op.yi, op.zi, lo, hi);
        if (ballot(op.active)) wv::stage<CAP>(g, cs, cell, lo, hi, R2, buf, op);
        if (in) row_ptr[i] = op.active ? op.count + 1 : 1;
        __syncthreads();
    }
}

// Fill pass: 16 lanes per row, four rows at a time.  For each staged batch
// and each group of four rows, lane p of a row tests staged candidates
// p, p + 16, ..., forms the FP64 weight of the ones it keeps, and the row's
// kept entries of each 16-candidate step are compacted by ballot into the
// row's next CSR slots: every store instruction writes four contiguous runs
// (one per row) instead of scattering one entry per lane over 64 rows, and
// the lanes stay busy when a cell holds fewer than 64 particles.  Entries keep
// the staged (ascending) order.  Dirichlet neighbours add to the diagonal but
// get no entry; the diagonal's value is written when the row is complete.
// Per-row state (position, next slot, diagonal sum and slot) lives in LDS.
constexpr int FILL_LANES = 16;

template <typename Real>
struct OpFill {
    const Real *x, *y, *z;
    const int* type;
    const unsigned char* bcv;
    Real tl;
    double re;
    int nrow;
    const Real *rx, *ry, *rz;
    const int* ri;
    const unsigned char* ract;
    int *rkpos, *rkdiag;
    double* rdiag;
    int* col;
    float* v32;
    double* v64;

    __device__ void load_cand(int j, wv::CandP<Real>& c)
    {
        c.x = x[j]; c.y = y[j]; c.z = z[j];
        c.j = j | (type[j] == 2 ? PPE_T2 : 0) | (bcv[j] ? PPE_BC : 0);
    }
    __device__ void process(const wv::CandP<Real>* buf, int cnt)
    {
        const int lane = threadIdx.x, p = lane % FILL_LANES, r4 = lane / FILL_LANES;
        const uint64_t seg = 0xffffull << (r4 * FILL_LANES);
        const uint64_t below = lane ? (~0ull >> (WAVE - lane)) : 0ull;
        for (int rg = 0; rg * (WAVE / FILL_LANES) < nrow; rg++) {
            const int r = rg * (WAVE / FILL_LANES) + r4;
            const bool rok = r < nrow && ract[r];
            Real xi = 0, yi = 0, zi = 0;
            int i = -1, kp = 0;
            if (rok) { xi = rx[r]; yi = ry[r]; zi = rz[r]; i = ri[r]; kp = rkpos[r]; }
            const double xd = (double)xi, yd = (double)yi, zd = (double)zi;
            double dsum = 0.0;
            int kd = -1;
            for (int k0 = 0; k0 < cnt; k0 += FILL_LANES) {
                const int k = k0 + p;
                const bool kin = k < cnt;
                const wv::CandP<Real> c = buf[kin ? k : 0];
                const Real rr = rr3(c.x - xi, c.y - yi, c.z - zi);
                const bool self = (c.j & PPE_JM) == i;
                const bool take = rok & kin & (self | ((rr < tl) & !(c.j & PPE_T2)));
                double w = 0.0;
                if (take && !self) {
                    const double dx = (double)c.x - xd, dy = (double)c.y - yd, dz = (double)c.z - zd;
                    w = __builtin_fma(re, rsq(dx * dx + dy * dy + dz * dz), -1.0);
                }
                dsum += w;
                const bool keep = take && (self || !(c.j & PPE_BC));   // Dirichlet neighbour: A_ij = 0
                const uint64_t m = ballot(keep) & seg;
                const int pos = kp + __popcll(m & below);
                if (keep) {
                    col[pos] = self ? i : (c.j & PPE_JM);
                    v64[pos] = w;
                    v32[pos] = (float)w;
                }
                kd = (keep && self) ? pos : kd;
                kp += __popcll(m);
            }
#pragma unroll
            for (int o = FILL_LANES / 2; o >= 1; o >>= 1) {
                dsum += __shfl_xor(dsum, o);
                const int t = __shfl_xor(kd, o);
                kd = t > kd ? t : kd;
            }
            if (rok && p == 0) {
                rkpos[r] = kp;
                rdiag[r] -= dsum;
                if (kd >= 0) rkdiag[r] = kd;
            }
        }
    }
};

template <typename Real, int CAP>
__global__ void __launch_bounds__(64) k_ppe_fill(Grid g, const int* __restrict__ cs, StateView<Real> s,
                                                 Thresh<Real> th, const unsigned char* __restrict__ bc,
                                                 const DevScalars* __restrict__ sc, double aii_over,
                                                 const int* __restrict__ row_ptr, int* __restrict__ col,
                                                 float* __restrict__ v32, double* __restrict__ v64)
{
    __shared__ wv::CandP<Real> buf[CAP];
    __shared__ Real rx[WAVE], ry[WAVE], rz[WAVE];
    __shared__ int ri[WAVE], rkpos[WAVE], rkdiag[WAVE];
    __shared__ double rdiag[WAVE];
    __shared__ unsigned char ract[WAVE];
    const int cell = xcd_cell(g.ncells);
    if (cell >= g.ncells) return;
    const int s0 = cs[cell], s1 = cs[cell + 1];
    // rho lambda alpha n0 / (2 d dt^2)
    const double dt = sc->dt;
    const double aii = aii_over * sc->lambda * sc->n0 / (dt * dt);
    OpFill<Real> op;
    op.x = s.x; op.y = s.y; op.z = s.z; op.type = s.type; op.bcv = bc;
    op.tl = th.t_lap; op.re = (double)th.re_lap;
    op.rx = rx; op.ry = ry; op.rz = rz; op.ri = ri; op.ract = ract;
    op.rkpos = rkpos; op.rkdiag = rkdiag; op.rdiag = rdiag;
    op.col = col; op.v32 = v32; op.v64 = v64;
    const Real R2 = wv::filt_r2(th.re_lap);
    const int lane = threadIdx.x;
    for (int c0 = s0; c0 < s1; c0 += WAVE) {
        const int i = c0 + lane;
        const bool in = i < s1;
        bool active = false;
        Real xi = 0, yi = 0, zi = 0;
        if (in) {
            xi = s.x[i]; yi = s.y[i]; zi = s.z[i];
            active = !bc[i] && s.type[i] != 2;
        }
        rx[lane] = xi; ry[lane] = yi; rz[lane] = zi; ri[lane] = i;
        ract[lane] = active ? 1 : 0;
        rkpos[lane] = in ? row_ptr[i] : 0;
        rkdiag[lane] = -1;
        rdiag[lane] = 0.0;
        op.nrow = s1 - c0 < WAVE ? s1 - c0 : WAVE;
        __syncthreads();
        Real lo[3], hi[3];
        wv::wave_box(active, xi, yi, zi, lo, hi);
        if (ballot(active)) wv::stage<CAP>(g, cs, cell, lo, hi, R2, buf, op);
        __syncthreads();
        if (in) {
            if (!active) {
                const int k = row_ptr[i];   // identity row
                col[k] = i; v64[k] = 1.0; v32[k] = 1.0f;
            } else if (rkdiag[lane] >= 0) {   // the row's own particle is always staged
                const double d = rdiag[lane] - aii;
                v64[rkdiag[lane]] = d;
                v32[rkdiag[lane]] = (float)d;
            }
        }
        __syncthreads();
    }
}

// ----------------------------------------------------------------- SpMV ---
// CSR-vector: G lanes per row, lane-strided partial sums, then a shuffle tree
// over the G lanes (fixed order: deterministic).  A lane loads its next U
// column/value pairs in one predicated batch (slots past the row end read
// nothing and add an exact zero), then issues the U gathers of p together, so
// a row of up to G*U entries costs two dependent memory latencies.  The lane
// still adds its entries in ascending order.
// batch depth of row_dot: with 16 lanes a row, 6 slots cover the ~93 entries
// of an re = 3.1 l0 row in one batch (measured best of 4/6/8; non-temporal
// matrix loads were slower)
constexpr int kU = 6;
template <int G, int U, typename MV, typename V, typename Acc>
__device__ __forceinline__ Acc row_dot(int row, int g, const int* __restrict__ rp,
                                       const int* __restrict__ col, const MV* __restrict__ val,
                                       const V* __restrict__ p)
{
    Acc a = (Acc)0;
    const int k1 = rp[row + 1];
    for (int k = rp[row] + g; k < k1; k += U * G) {
        int c[U];
        MV v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int kk = k + u * G;
            c[u] = kk < k1 ? col[kk] : -1;
            v[u] = kk < k1 ? val[kk] : (MV)0;
        }
        V pv[U];
#pragma unroll
        for (int u = 0; u < U; u++) pv[u] = c[u] >= 0 ? p[c[u]] : (V)0;
#pragma unroll
        for (int u = 0; u < U; u++) a += (Acc)v[u] * (Acc)pv[u];
    }
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1) a += __shfl_xor(a, m);
    return a;
}

// The inner CG runs on a fixed grid (kCgBlocks x kCgThreads, grid-stride
// loops).  A kernel that produces a dot product writes one partial per block;
// every block of the next kernel sums all partials itself, in the same fixed
// order, so each block holds the identical value without atomics or a wait on
// other blocks.  Block 0 publishes the scalars the host reads.
constexpr int kCgBlocks = 1024;
constexpr int kCgThreads = 512;

__device__ __forceinline__ void block_partial(double v, double* __restrict__ partial)
{
    __shared__ double red[kCgThreads / WAVE];
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    v = wave_sum(v);
    if (lane == 0) red[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < kCgThreads / WAVE; k++) s += red[k];
        partial[blockIdx.x] = s;
    }
    __syncthreads();   // red[] is free again for a following call
}

__device__ __forceinline__ double block_total(const double* __restrict__ partial, int np)
{
    __shared__ double red[kCgThreads / WAVE];
    __shared__ double tot;
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    double s = 0.0;
    for (int k = threadIdx.x; k < np; k += kCgThreads) s += partial[k];
    s = wave_sum(s);
    if (lane == 0) red[w] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < kCgThreads / WAVE; k++) t += red[k];
        tot = t;
    }
    __syncthreads();
    return tot;
}

// q = A p; partial sums of p . q
template <int G, typename MV, typename V>
__global__ void __launch_bounds__(kCgThreads) k_spmv_dot(int n, const int* __restrict__ rp,
                                                         const int* __restrict__ col, const MV* __restrict__ val,
                                                         const V* __restrict__ p, V* __restrict__ q,
                                                         double* __restrict__ part_pq, const PpeScalars* ps)
{
    if (ps->done) return;
    const int g = threadIdx.x % G, rpb = kCgThreads / G;
    double d = 0.0;
    for (int row = blockIdx.x * rpb + threadIdx.x / G; row < n; row += gridDim.x * rpb) {
        const V a = row_dot<G, kU, MV, V, V>(row, g, rp, col, val, p);
        if (g == 0) {
            q[row] = a;
            d += (double)p[row] * (double)a;
        }
    }
    block_partial(d, part_pq);
}

// pq = sum of the SpMV partials; x += alpha p, r -= alpha q with alpha = rr / pq;
// partial sums of r . r.  The MPS Poisson matrix is negative definite: p.Ap < 0
// is normal; only a vanishing or non-finite p.Ap is a breakdown.
template <typename V>
__global__ void __launch_bounds__(kCgThreads) k_cg_update(int n, V* __restrict__ x, V* __restrict__ r,
                                                          const V* __restrict__ p, const V* __restrict__ q,
                                                          const double* __restrict__ part_pq,
                                                          double* __restrict__ part_rr, PpeScalars* ps)
{
    if (ps->done) return;
    const double pq = block_total(part_pq, gridDim.x);
    const double rr = ps->rr;
    if (!(fabs(pq) > 0.0 && isfinite(pq))) {
        if (blockIdx.x == 0 && threadIdx.x == 0) ps->done = 2;
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ps->pq = pq;
        ps->rr_old = rr;   // k_cg_dir rewrites rr while its blocks still need this one
    }
    const V alpha = (V)(rr / pq);
    double d = 0.0;
    for (int i = blockIdx.x * kCgThreads + threadIdx.x; i < n; i += gridDim.x * kCgThreads) {
        x[i] += alpha * p[i];
        const V ri = r[i] - alpha * q[i];
        r[i] = ri;
        d += (double)ri * (double)ri;
    }
    block_partial(d, part_rr);
}

// rr_new = sum of the update partials; beta = rr_new / rr; p = r + beta p.
// Block 0 publishes rr, beta, the iteration count and the stop flag.
template <typename V>
__global__ void __launch_bounds__(kCgThreads) k_cg_dir(int n, const V* __restrict__ r, V* __restrict__ p,
                                                       const double* __restrict__ part_rr, PpeScalars* ps)
{
    if (ps->done) return;
    const double tot = block_total(part_rr, gridDim.x);
    const double rr = ps->rr_old;
    const double beta = tot / rr;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ps->beta = beta;
        ps->rr = tot;
        ps->iters += 1;
        if (!isfinite(tot)) ps->done = 2;
        else if (!(tot > ps->tol2)) ps->done = 1;
    }
    const V bv = (V)beta;
    for (int i = blockIdx.x * kCgThreads + threadIdx.x; i < n; i += gridDim.x * kCgThreads)
        p[i] = r[i] + bv * p[i];
}

// Outer residual in FP64: r64 = b - A64 x; rr64 = r64 . r64.  Also starts the
// inner solve A e = r: e = 0, r_in = p = r64 (in the inner precision),
// rr = r_in . r_in, tol2 = max(tol_in^2 rr, floor2), done = 0.
template <int G, typename V>
__global__ void __launch_bounds__(kCgThreads) k_residual(int n, const int* __restrict__ rp, const int* __restrict__ col,
                                                         const double* __restrict__ val, const double* __restrict__ b,
                                                         const double* __restrict__ x, V* __restrict__ e,
                                                         V* __restrict__ r, V* __restrict__ p,
                                                         double* __restrict__ part64, double* __restrict__ part_in)
{
    const int g = threadIdx.x % G, rpb = kCgThreads / G;
    double d64 = 0.0, din = 0.0;
    for (int row = blockIdx.x * rpb + threadIdx.x / G; row < n; row += gridDim.x * rpb) {
        const double a = row_dot<G, kU, double, double, double>(row, g, rp, col, val, x);
        if (g == 0) {
            const double ri = b[row] - a;
            const V rv = (V)ri;
            e[row] = (V)0;
            r[row] = rv;
            p[row] = rv;
            d64 += ri * ri;
            din += (double)rv * (double)rv;
        }
    }
    block_partial(d64, part64);
    block_partial(din, part_in);
}

// one block: totals of the residual partials and the start of the inner solve
__global__ void __launch_bounds__(kCgThreads) k_residual_fin(const double* __restrict__ part64,
                                                             const double* __restrict__ part_in, int np,
                                                             PpeScalars* ps, double tol_in2, double floor2)
{
    const double tot = block_total(part64, np);
    const double rin = block_total(part_in, np);
    if (threadIdx.x == 0) {
        ps->rr64 = tot;
        ps->rin = rin;
        ps->rr = rin;
        ps->rr0 = rin;
        ps->tol2 = fmax(tol_in2 * rin, floor2);
        ps->done = rin > ps->tol2 ? 0 : 1;
        ps->iters = 0;
    }
}

template <typename V>
__global__ void __launch_bounds__(256) k_ir_update(int n, double* __restrict__ x, const V* __restrict__ e)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += (double)e[i];
}

// pressCorr (MPS_CPU.cpp:734-743): p = max(x, 0), written into the state.
template <typename Real>
__global__ void __launch_bounds__(256) k_press_corr(int n, const double* __restrict__ x, Real* __restrict__ p)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (Real)(x[i] > 0.0 ? x[i] : 0.0);
}

// ------------------------------------------------------------- launchers ---
static inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }
constexpr int kG = 16;
// Reducing kernels run at most kRedBlocks blocks (grid-stride loops): the
// last-block reduction then takes one ticket atomic per block, a few thousand
// per launch instead of one per 16 rows.  The grid depends on n only, so the
// summation order, and the result, is fixed for a given system.
constexpr int kRedBlocks = 2048;
static inline int red_grid(long threads) { return std::min(cdiv(threads, 256), kRedBlocks); }

// doubles in each partial buffer: the block partials of the reducing kernels,
// and also the uint32 tile totals of the row-pointer scan (4096 rows a tile)
int ppe_partial_blocks(int n) { return std::max(std::max(kRedBlocks, kCgBlocks), cdiv((long)n + 1, 2 * 4096)) + 8; }

template <typename Real>
void launch_ppe_rows(int n, StateView<Real> s, const DevScalars* sc, double beta, double btmp_over,
                     int x0_kind, const double* x0_ext, unsigned char* bc, double* b, double* x,
                     double* partial, PpeScalars* ps, hipStream_t st)
{
    hipLaunchKernelGGL(k_ppe_rows<Real>, dim3(red_grid(n)), dim3(256), 0, st, n, s, sc, beta,
                       btmp_over, x0_kind, x0_ext, bc, b, x, partial, ps);
}

template <typename Real>
void launch_ppe_assemble(const Grid& g, const int* cs, StateView<Real> s, const Thresh<Real>& th,
                         const unsigned char* bc, const DevScalars* sc, double aii_over, int fill,
                         int* row_ptr, int* col, float* v32, double* v64, hipStream_t st)
{
    if (!fill) {
        hipLaunchKernelGGL((k_ppe_count<Real, wv::cap_of<Real>(128)>), dim3(xcd_grid(g.ncells)), dim3(WAVE), 0, st,
                           g, cs, s, th, bc, row_ptr);
        return;
    }
    hipLaunchKernelGGL((k_ppe_fill<Real, wv::cap_of<Real>(128)>), dim3(xcd_grid(g.ncells)), dim3(WAVE), 0, st, g,
                       cs, s, th, bc, sc, aii_over, row_ptr, col, v32, v64);
}

template <typename MV, typename V>
void launch_cg_iter(int n, const int* rp, const int* col, const MV* val, V* x, V* r, V* p, V* q,
                    double* partial, double* partial2, PpeScalars* ps, hipStream_t st)
{
    hipLaunchKernelGGL((k_spmv_dot<kG, MV, V>), dim3(kCgBlocks), dim3(kCgThreads), 0, st, n, rp, col, val, p, q,
                       partial, ps);
    hipLaunchKernelGGL(k_cg_update<V>, dim3(kCgBlocks), dim3(kCgThreads), 0, st, n, x, r, p, q, partial,
                       partial2, ps);
    hipLaunchKernelGGL(k_cg_dir<V>, dim3(kCgBlocks), dim3(kCgThreads), 0, st, n, r, p, partial2, ps);
}

template <typename V>
void launch_residual(int n, const int* rp, const int* col, const double* val, const double* b,
                     const double* x, V* e, V* r, V* p, double* partial, double* partial2,
                     PpeScalars* ps, double tol_in2, double floor2, hipStream_t st)
{
    hipLaunchKernelGGL((k_residual<kG, V>), dim3(kCgBlocks), dim3(kCgThreads), 0, st, n, rp, col, val, b, x, e, r,
                       p, partial, partial2);
    hipLaunchKernelGGL(k_residual_fin, dim3(1), dim3(kCgThreads), 0, st, partial, partial2, kCgBlocks, ps, tol_in2,
                       floor2);
}

template <typename V>
void launch_ir_update(int n, double* x, const V* e, hipStream_t st)
{
    hipLaunchKernelGGL(k_ir_update<V>, dim3(cdiv(n, 256)), dim3(256), 0, st, n, x, e);
}

template <typename Real>
void launch_press_corr(int n, const double* x, Real* p, hipStream_t st)
{
    hipLaunchKernelGGL(k_press_corr<Real>, dim3(cdiv(n, 256)), dim3(256), 0, st, n, x, p);
}

template void launch_ppe_rows<float>(int, StateView<float>, const DevScalars*, double, double, int,
                                     const double*, unsigned char*, double*, double*, double*,
                                     PpeScalars*, hipStream_t);
template void launch_ppe_rows<double>(int, StateView<double>, const DevScalars*, double, double, int,
                                      const double*, unsigned char*, double*, double*, double*,
                                      PpeScalars*, hipStream_t);
template void launch_ppe_assemble<float>(const Grid&, const int*, StateView<float>, const Thresh<float>&,
                                         const unsigned char*, const DevScalars*, double, int, int*,
                                         int*, float*, double*, hipStream_t);
template void launch_ppe_assemble<double>(const Grid&, const int*, StateView<double>,
                                          const Thresh<double>&, const unsigned char*,
                                          const DevScalars*, double, int, int*, int*, float*, double*,
                                          hipStream_t);
template void launch_cg_iter<float, float>(int, const int*, const int*, const float*, float*, float*,
                                           float*, float*, double*, double*, PpeScalars*, hipStream_t);
template void launch_cg_iter<double, double>(int, const int*, const int*, const double*, double*,
                                             double*, double*, double*, double*, double*, PpeScalars*,
                                             hipStream_t);
template void launch_residual<float>(int, const int*, const int*, const double*, const double*,
                                     const double*, float*, float*, float*, double*, double*,
                                     PpeScalars*, double, double, hipStream_t);
template void launch_residual<double>(int, const int*, const int*, const double*, const double*,
                                      const double*, double*, double*, double*, double*, double*,
                                      PpeScalars*, double, double, hipStream_t);
template void launch_ir_update<float>(int, double*, const float*, hipStream_t);
template void launch_ir_update<double>(int, double*, const double*, hipStream_t);
template void launch_press_corr<float>(int, const double*, float*, hipStream_t);
template void launch_press_corr<double>(int, const double*, double*, hipStream_t);

}  // namespace mps
