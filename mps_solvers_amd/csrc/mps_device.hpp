// mps_device.hpp -- shared device-side helpers for the gfx950 MPS kernels.
//
// Layout conventions (HBM, structure of arrays, cell-sorted):
//   x, y, z, u, v, w, p : Real (float on the mixed-precision path, double on the FP64 path)
//   n                   : double (particle number density, always FP64)
//   type, id, normal    : int32
//   cell_start[ncells+1]: CSR offsets of the cell-sorted order
// Every neighbour sweep runs one 64-lane wavefront per cell (chunks of 64
// particles), stages the 27-cell candidate set through LDS after a
// wave-level bounding-box filter, and broadcasts each staged candidate to
// all lanes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mps {

constexpr int WAVE = 64;

struct Grid {
    double left, back, bottom, h;   // cell origin and cell size (exact values of the state precision)
    int dx, dy, dz, sheet, ncells, pad_;
};

// Per-precision thresholds.  The neighbour tests are taken in the state
// precision exactly as the reference takes them:
//   w(r; re) != 0  <=>  sqrt(rr) < re  <=>  rr < t_*    (t_* precomputed on the host)
//   collision hit   <=>  sqrt(rr) <= col <=>  rr < t_col
//   p-hat candidate <=>  rr <= re_n^2 (re_n^2 rounded in the state precision)
template <typename Real>
struct Thresh {
    Real re_n, re_lap, col, re_n2;
    Real t_n, t_lap, t_col;
    Real pad_;
};

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ int lanes_below(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ int lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

template <typename T>
__device__ __forceinline__ T wave_min(T v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) { T o = __shfl_xor(v, m); v = o < v ? o : v; }
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) { T o = __shfl_xor(v, m); v = o > v ? o : v; }
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
    return v;
}

// Squared distance formed exactly as the reference's real3 dot product:
// ((dx*dx + dy*dy) + dz*dz), no contraction (file built with -ffp-contract=off).
template <typename Real>
__device__ __forceinline__ Real rr3(Real dx, Real dy, Real dz)
{
    Real a = dx * dx;
    Real b = dy * dy;
    Real c = dz * dz;
    Real s = a + b;
    return s + c;
}

// Fused multiply-add in the operands' own precision (the bare __builtin_fma is
// the double one: float operands would be widened, fused in FP64 and rounded
// back, at several times the VALU cost).
__device__ __forceinline__ float fmar(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmar(double a, double b, double c) { return __builtin_fma(a, b, c); }

__device__ __forceinline__ float rsq(float a) { return __builtin_amdgcn_rsqf(a); }
__device__ __forceinline__ double rsq(double a)
{
    double y = __builtin_amdgcn_rsq(a);
    // one Newton step on the hardware estimate
    double h = 0.5 * a * y * y;
    return y * (1.5 - h);
}

// Map a block to a cell so that each of the 8 XCDs walks a contiguous run of
// cells (blocks are dealt round-robin to XCDs): neighbouring cells then
// share the XCD's L2.
__device__ __forceinline__ int xcd_cell(int ncells)
{
    int per = (int)(gridDim.x >> 3);
    int b = (int)blockIdx.x;
    return (b & 7) * per + (b >> 3);
}
inline int xcd_grid(int ncells) { return ((ncells + 7) / 8) * 8; }

__device__ __forceinline__ void cell_xyz(const Grid& g, int c, int& cx, int& cy, int& cz)
{
    cz = c / g.sheet;
    int r = c - cz * g.sheet;
    cy = r / g.dx;
    cx = r - cy * g.dx;
}

}  // namespace mps
