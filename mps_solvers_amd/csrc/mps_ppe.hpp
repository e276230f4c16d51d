// mps_ppe.hpp -- pressure Poisson solver: device scalars, deterministic grid
// reductions and launchers.
#pragma once

#include <hip/hip_runtime.h>

#include "mps_device.hpp"
#include "mps_kernels.hpp"

namespace mps {

struct PpeScalars {
    double rr, rr0, pq, beta, tol2;   // inner CG
    double rr_old;                    // rr of the previous inner iteration
    double rr64, bb, rin;             // outer residual, |b|^2, inner start
    int done, iters;                  // done: 1 converged, 2 breakdown (p.Ap <= 0)
    unsigned int ticket[4];
};

// Sum v over the grid.  Every block stores its partial; the last block to
// arrive sums the partials in a fixed order (deterministic), writes *out from
// thread 0 and resets the ticket.  Returns true in the last block.
__device__ __forceinline__ bool block_sum_last2(double v, double v2, double* partial, double* partial2,
                                                unsigned int* ticket, double* out, double* out2)
{
    __shared__ double red[2][8];
    __shared__ int last;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    v = wave_sum(v);
    v2 = wave_sum(v2);
    if (lane == 0) { red[0][w] = v; red[1][w] = v2; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0, s2 = 0.0;
        for (int k = 0; k < nw; k++) { s += red[0][k]; s2 += red[1][k]; }
        partial[blockIdx.x] = s;
        if (partial2) partial2[blockIdx.x] = s2;
        __threadfence();
        const unsigned int t = atomicAdd(ticket, 1u);
        last = t == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    double s = 0.0, s2 = 0.0;
    for (int k = threadIdx.x; k < (int)gridDim.x; k += blockDim.x) {
        s += ((volatile double*)partial)[k];
        if (partial2) s2 += ((volatile double*)partial2)[k];
    }
    s = wave_sum(s);
    s2 = wave_sum(s2);
    __syncthreads();
    if (lane == 0) { red[0][w] = s; red[1][w] = s2; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0, t2 = 0.0;
        for (int k = 0; k < nw; k++) { t += red[0][k]; t2 += red[1][k]; }
        *out = t;
        if (out2) *out2 = t2;
        *ticket = 0u;
    }
    return true;
}

__device__ __forceinline__ bool block_sum_last(double v, double* partial, unsigned int* ticket, double* out)
{
    return block_sum_last2(v, 0.0, partial, nullptr, ticket, out, nullptr);
}

int ppe_partial_blocks(int n);

template <typename Real>
void launch_ppe_rows(int n, StateView<Real> s, const DevScalars* sc, double beta, double btmp_over,
                     int x0_kind, const double* x0_ext, unsigned char* bc, double* b, double* x,
                     double* partial, PpeScalars* ps, hipStream_t st);
template <typename Real>
void launch_ppe_assemble(const Grid& g, const int* cs, StateView<Real> s, const Thresh<Real>& th,
                         const unsigned char* bc, const DevScalars* sc, double aii_over, int fill,
                         int* row_ptr, int* col, float* v32, double* v64, hipStream_t st);
template <typename MV, typename V>
void launch_cg_iter(int n, const int* rp, const int* col, const MV* val, V* x, V* r, V* p, V* q,
                    double* partial, double* partial2, PpeScalars* ps, hipStream_t st);
template <typename V>
void launch_residual(int n, const int* rp, const int* col, const double* val, const double* b,
                     const double* x, V* e, V* r, V* p, double* partial, double* partial2,
                     PpeScalars* ps, double tol_in2, double floor2, hipStream_t st);
template <typename V>
void launch_ir_update(int n, double* x, const V* e, hipStream_t st);
template <typename Real>
void launch_press_corr(int n, const double* x, Real* p, hipStream_t st);
void launch_scan_excl(uint32_t* a, int m, uint32_t* tiles, hipStream_t st);

}  // namespace mps
