// mps_stage_wave.hpp -- one-wavefront-per-cell LDS staging of the 27-cell
// candidate set, shared by the neighbour sweeps and the Poisson assembly.
#pragma once

#include <float.h>

#include "mps_device.hpp"

namespace mps {
namespace wv {

// Staged candidate of the sweeps that need velocity or pressure (32 B for float).
template <typename Real>
struct alignas(16) Cand {
    Real x, y, z;
    int j, t;
    Real a, b, c;
};

// Staged candidate of the number-density sweep: position and index only
// (16 B for float: one LDS read per broadcast).
template <typename Real>
struct alignas(16) CandP {
    Real x, y, z;
    int j;
};

// LDS batch capacity in candidates for a float capacity `cap`: FP64
// candidates are twice as large, so half as many.
template <typename Real> __host__ __device__ constexpr int cap_of(int cap) { return sizeof(Real) == 4 ? cap : cap / 2; }

template <typename Real> __device__ __forceinline__ Real big();
template <> __device__ __forceinline__ float big<float>() { return FLT_MAX; }
template <> __device__ __forceinline__ double big<double>() { return DBL_MAX; }

template <typename Real>
__device__ __forceinline__ Real box_d2(Real qx, Real qy, Real qz, const Real* lo, const Real* hi)
{
    Real dx = fmax(fmax(lo[0] - qx, qx - hi[0]), (Real)0);
    Real dy = fmax(fmax(lo[1] - qy, qy - hi[1]), (Real)0);
    Real dz = fmax(fmax(lo[2] - qz, qz - hi[2]), (Real)0);
    return dx * dx + dy * dy + dz * dz;
}

struct StageResult { int count; bool overflowed; };

// Distance along one axis from the box [lo, hi] to the slab [a, b] (0 if they overlap).
template <typename Real>
__device__ __forceinline__ Real axis_gap(Real lo, Real hi, Real a, Real b)
{
    return fmax(fmax(a - hi, lo - b), (Real)0);
}

// Visit buf[0..cnt) in ascending order.  Candidates are read from LDS U at a
// time before the first of them is visited, so that the LDS latency of a
// group overlaps the arithmetic of the group before it (a plain loop waits
// on every read).  U trades that overlap against registers: each sweep uses
// the U measured best for it (a larger U that costs a resident wavefront per
// SIMD loses more than it hides).
template <int U, class Cd, class F>
__device__ __forceinline__ void for_each_cand(const Cd* buf, int cnt, F&& f)
{
    int k = 0;
    for (; k + U <= cnt; k += U) {
        Cd c[U];
#pragma unroll
        for (int u = 0; u < U; u++) c[u] = buf[k + u];
#pragma unroll
        for (int u = 0; u < U; u++) f(c[u], k + u);
    }
    for (; k < cnt; k++) f(buf[k], k);
}

// Stream the stencil of `cell`, keep candidates within sqrt(R2) of the box,
// and hand full LDS batches (CAP candidates of type Cd) to op.process().  The
// last batch stays in `buf`.  With CULL, stencil cells whose extent (widened
// by 1e-4 h, far beyond the rounding of the hash) lies at least sqrt(R2) from
// the box are not read: none of their particles could pass the box filter.
// Only the collision radius (0.8 l0 against 4 l0 cells) leaves such cells;
// the larger radii would pay the extra per-row arithmetic for nothing.
template <int CAP, bool CULL = false, typename Real, class Op, class Cd>
__device__ StageResult stage(const Grid& g, const int* __restrict__ cs, int cell, const Real* lo,
                             const Real* hi, Real R2, Cd* buf, Op& op)
{
    const int lane = threadIdx.x;
    int cx, cy, cz;
    cell_xyz(g, cell, cx, cy, cz);
    const Real h = (Real)g.h, mg = (Real)1e-4 * h;
    const Real left = (Real)g.left, back = (Real)g.back, bottom = (Real)g.bottom;
    int x0 = cx > 0 ? cx - 1 : 0, x1 = cx + 1 < g.dx ? cx + 1 : g.dx - 1;
    if (CULL) {
        const Real gl = axis_gap(lo[0], hi[0], left + (Real)x0 * h - mg, left + (Real)(x0 + 1) * h + mg);
        const Real gr = axis_gap(lo[0], hi[0], left + (Real)x1 * h - mg, left + (Real)(x1 + 1) * h + mg);
        // the x-end cells of every stencil row are at least this far away
        if (x0 < cx && gl * gl >= R2) x0 = cx;
        if (x1 > cx && gr * gr >= R2) x1 = cx;
    }
    int count = 0;
    bool over = false;
    for (int dz = -1; dz <= 1; dz++) {
        const int zz = cz + dz;
        if (zz < 0 || zz >= g.dz) continue;
        const Real gz = CULL ? axis_gap(lo[2], hi[2], bottom + (Real)zz * h - mg, bottom + (Real)(zz + 1) * h + mg)
                             : (Real)0;
        if (CULL && gz * gz >= R2) continue;
        for (int dy = -1; dy <= 1; dy++) {
            const int yy = cy + dy;
            if (yy < 0 || yy >= g.dy) continue;
            int r0 = x0, r1 = x1;
            if (CULL) {
                const Real gy = axis_gap(lo[1], hi[1], back + (Real)yy * h - mg, back + (Real)(yy + 1) * h + mg);
                const Real gyz = gz * gz + gy * gy;
                if (gyz >= R2) continue;
                if (r0 < cx) {
                    const Real gx = axis_gap(lo[0], hi[0], left + (Real)r0 * h - mg, left + (Real)(r0 + 1) * h + mg);
                    if (gyz + gx * gx >= R2) r0 = cx;
                }
                if (r1 > cx) {
                    const Real gx = axis_gap(lo[0], hi[0], left + (Real)r1 * h - mg, left + (Real)(r1 + 1) * h + mg);
                    if (gyz + gx * gx >= R2) r1 = cx;
                }
            }
            const int row = zz * g.sheet + yy * g.dx;
            const int a = cs[row + r0], b = cs[row + r1 + 1];
            for (int base = a; base < b; base += WAVE) {
                const int j = base + lane;
                Cd c;
                bool keep = false;
                if (j < b) {
                    op.load_cand(j, c);
                    keep = box_d2(c.x, c.y, c.z, lo, hi) < R2;
                }
                const uint64_t m = ballot(keep);
                if (keep) buf[count + lanes_below(m)] = c;
                count += __popcll(m);
                if (count > CAP - WAVE) {
                    __syncthreads();
                    op.process(buf, count);
                    __syncthreads();
                    count = 0;
                    over = true;
                }
            }
        }
    }
    __syncthreads();
    if (count) op.process(buf, count);
    return {count, over};
}

// Bounding box of the active lanes, widened by one part in 1e5 so that the
// staging filter never drops a candidate the exact tests would keep.
template <typename Real>
__device__ __forceinline__ void wave_box(bool act, Real x, Real y, Real z, Real* lo, Real* hi)
{
    lo[0] = wave_min(act ? x : big<Real>());
    lo[1] = wave_min(act ? y : big<Real>());
    lo[2] = wave_min(act ? z : big<Real>());
    hi[0] = wave_max(act ? x : -big<Real>());
    hi[1] = wave_max(act ? y : -big<Real>());
    hi[2] = wave_max(act ? z : -big<Real>());
}

template <typename Real>
__device__ __forceinline__ Real filt_r2(Real R) { Real r = R * (Real)1.00001 + (Real)1e-12; return r * r; }

// ------------------------------------------------------------ lane groups ---
// A chunk of up to 64 particles is cut into ng <= G groups of at most
// S = 64 / G particles, and each group stages and walks its own candidate set
// with G lanes per particle: lane q*S + s takes the staged candidates
// q, q + G, q + 2G, ... (phase q) for the group's s-th particle, and the G
// partial sums are combined at the end.  A group's box is a fraction of the
// chunk's, so each particle evaluates fewer candidates; the price is one
// staging pass per group and the final combine.

// Far-away candidate used to pad a batch to a multiple of G: every cutoff
// test fails on it and it carries zero values, so it adds exactly nothing.
template <typename Real>
__device__ __forceinline__ void far_cand(Cand<Real>& c)
{
    c.x = c.y = c.z = (Real)1e30;
    c.j = -1;
    c.t = 0;
    c.a = c.b = c.c = (Real)0;
}

// Pad buf[0..cnt) with far candidates to a multiple of G; returns the
// candidates per phase.  buf must hold cnt + G - 1 entries.
template <int G, class Cd>
__device__ __forceinline__ int pad_phases(Cd* buf, int cnt)
{
    if (G == 1) return cnt;
    const int padded = (cnt + G - 1) / G * G;
    if ((int)threadIdx.x < padded - cnt) {
        Cd c;
        far_cand(c);
        buf[cnt + threadIdx.x] = c;
    }
    __syncthreads();
    return padded / G;
}

// Visit phase q's candidates buf[k*G + q], k in [k0, k1), in ascending order,
// reading U of them from LDS before the first is visited (as for_each_cand).
template <int U, int G, class Cd, class F>
__device__ __forceinline__ void for_each_phase(const Cd* buf, int k0, int k1, int q, F&& f)
{
    int k = k0;
    for (; k + U <= k1; k += U) {
        Cd c[U];
#pragma unroll
        for (int u = 0; u < U; u++) c[u] = buf[(k + u) * G + q];
#pragma unroll
        for (int u = 0; u < U; u++) f(c[u], (k + u) * G + q);
    }
    for (; k < k1; k++) f(buf[k * G + q], k * G + q);
}

// Combine the G phase lanes of each particle (lanes s, s + S, s + 2S, ...).
template <int G, typename T>
__device__ __forceinline__ T phase_sum(T v)
{
#pragma unroll
    for (int m = WAVE / G; m < WAVE; m <<= 1) v += __shfl_xor(v, m);
    return v;
}

// Cut the active lanes into groups: ng = the smallest power of two <= G with
// ng * S >= (active lanes); each level halves every group at the median of a
// 6-bit quantised coordinate along the longest remaining axis of the chunk's
// box (ties by lane), so group sizes differ by at most one and none exceeds S.
// On return ord[g * S + s] is the lane of group g's s-th particle (ascending
// lanes within a group) or -1.  ord holds 64 ints.
template <int G, typename Real>
__device__ int split_groups(bool act, Real x, Real y, Real z, const Real* lo, const Real* hi, int* ord)
{
    constexpr int S = WAVE / G;
    const int lane = threadIdx.x;
    const uint64_t am = ballot(act);
    const int n = __popcll(am);
    int ng = 1;
    while (ng < G && ng * S < n) ng <<= 1;
    ord[lane] = -1;
    const uint64_t below = lane ? (~0ull >> (WAVE - lane)) : 0ull;
    uint64_t mine = am;
    int grp = 0;
    Real ext[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    for (int lv = 1; lv < ng; lv <<= 1) {
        int a = ext[1] > ext[0] ? 1 : 0;
        a = ext[2] > ext[a] ? 2 : a;
        const Real e = ext[a];
        ext[a] = e * (Real)0.5;
        const Real c = a == 0 ? x : (a == 1 ? y : z);
        int key = 0;
        if (e > (Real)0) key = min(63, max(0, (int)((c - lo[a]) * ((Real)64 / e))));
        uint64_t eq = mine;
        int less = 0;
#pragma unroll
        for (int b = 5; b >= 0; b--) {
            const bool bit = (key >> b) & 1;
            const uint64_t m1 = ballot(bit);
            if (bit) less += __popcll(eq & ~m1);
            eq &= bit ? m1 : ~m1;
        }
        less += __popcll(eq & below);
        const bool up = less >= (__popcll(mine) >> 1);
        const uint64_t mu = ballot(act && up);
        mine &= up ? mu : ~mu;
        grp = 2 * grp + (up ? 1 : 0);
    }
    __syncthreads();
    if (act) ord[grp * S + __popcll(mine & below)] = lane;
    __syncthreads();
    return ng;
}

}  // namespace wv
}  // namespace mps
