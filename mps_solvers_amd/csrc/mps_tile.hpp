// mps_tile.hpp -- two-level LDS staging for the neighbour sweeps on gfx950.
//
// One 512-thread block (8 wavefronts) per cell, in rounds of up to 64
// particles.  Wave w owns a tile of 8 consecutive particles; each particle is
// served by 8 lanes that split its candidate list (lane = 8 * particle + sub).
//
//   1. The block streams the 27-cell stencil (9 contiguous rows of the
//      cell-sorted arrays) with coalesced loads, keeps candidates within R of
//      the bounding box of the round's active particles, and compacts them in
//      ascending index order into LDS (array S).
//   2. Each wave filters S again against the box of its own 8 particles into a
//      private u16 index list L[w] (ascending).
//   3. The 8 lanes of a particle walk L[w] with stride 8; partial results are
//      combined with a 3-step shuffle tree at the end.
//
// The tile of 8 particles sees about half the candidates the whole cell does
// (Minkowski sum of a ~3x1 l0 tile with the cutoff sphere instead of a
// 3x3x3 l0 cell), which halves the VALU work per particle; the exact
// within-cutoff tests are unchanged.
#pragma once

#include <float.h>

#include "mps_device.hpp"

namespace mps {

constexpr int TB = 512;   // threads per block
constexpr int TW = 8;     // waves per block
constexpr int TK = 8;     // lanes per particle

template <typename Real>
struct alignas(16) Cand {
    Real x, y, z;
    int j, t;
    Real a, b, c;
};

template <typename Real> struct Cap2 { static constexpr int v = 1024; };
template <> struct Cap2<double> { static constexpr int v = 512; };

template <typename Real>
struct TileShared {
    Cand<Real> S[Cap2<Real>::v];
    uint16_t L[TW][Cap2<Real>::v];
    Real box[TW][6];
    int wcnt[TW];
};

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

// widened by one part in 1e5 so that the filters never drop a candidate the
// exact tests would keep
template <typename Real>
__device__ __forceinline__ Real filt_r2(Real R) { Real r = R * (Real)1.00001 + (Real)1e-12; return r * r; }

template <typename T>
__device__ __forceinline__ T sub_sum(T v)
{
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 4);
    return v;
}
template <typename T>
__device__ __forceinline__ T sub_min(T v)
{
    T o = __shfl_xor(v, 1); v = o < v ? o : v;
    o = __shfl_xor(v, 2); v = o < v ? o : v;
    o = __shfl_xor(v, 4); v = o < v ? o : v;
    return v;
}
template <typename T>
__device__ __forceinline__ T sub_max(T v)
{
    T o = __shfl_xor(v, 1); v = o > v ? o : v;
    o = __shfl_xor(v, 2); v = o > v ? o : v;
    o = __shfl_xor(v, 4); v = o > v ? o : v;
    return v;
}

__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per-round geometry: the wave's tile box (uniform in the wave) and the
// block box over all active tiles.  Returns false when no tile is active.
template <typename Real>
__device__ __forceinline__ bool tile_boxes(TileShared<Real>& sh, bool act, Real x, Real y, Real z,
                                           Real* wlo, Real* whi, Real* blo, Real* bhi, bool& tile_act)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    wlo[0] = wave_min(act ? x : big<Real>());
    wlo[1] = wave_min(act ? y : big<Real>());
    wlo[2] = wave_min(act ? z : big<Real>());
    whi[0] = wave_max(act ? x : -big<Real>());
    whi[1] = wave_max(act ? y : -big<Real>());
    whi[2] = wave_max(act ? z : -big<Real>());
    tile_act = ballot(act) != 0ull;
    if (lane == 0) {
        sh.box[wave][0] = wlo[0]; sh.box[wave][1] = wlo[1]; sh.box[wave][2] = wlo[2];
        sh.box[wave][3] = whi[0]; sh.box[wave][4] = whi[1]; sh.box[wave][5] = whi[2];
    }
    __syncthreads();
    blo[0] = blo[1] = blo[2] = big<Real>();
    bhi[0] = bhi[1] = bhi[2] = -big<Real>();
    for (int w = 0; w < TW; w++)
        for (int a = 0; a < 3; a++) {
            blo[a] = fmin(blo[a], sh.box[w][a]);
            bhi[a] = fmax(bhi[a], sh.box[w][3 + a]);
        }
    __syncthreads();
    return blo[0] <= bhi[0];
}

// Filter S[0, count) against the wave's tile box into L[wave] and run
// op.process over it.  Ends with a block barrier (S and L reusable).
template <typename Real, class Op>
__device__ __forceinline__ int tile_batch(TileShared<Real>& sh, int count, const Real* wlo, const Real* whi,
                                          Real R2, bool tile_act, Op& op)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int len = 0;
    if (tile_act) {
        for (int k0 = 0; k0 < count; k0 += WAVE) {
            const int k = k0 + lane;
            bool keep = false;
            if (k < count) keep = box_d2(sh.S[k].x, sh.S[k].y, sh.S[k].z, wlo, whi) < R2;
            const uint64_t m = ballot(keep);
            if (keep) sh.L[wave][len + lanes_below(m)] = (uint16_t)k;
            len += __popcll(m);
        }
        wave_sync_lds();
        op.process(sh.S, sh.L[wave], len);
    }
    __syncthreads();
    return len;
}

struct TileStage { int count, len; bool overflowed; };

// Stream the stencil of `cell` through S in batches; each full batch is
// filtered and processed.  The last batch stays in S / L (see TileStage).
template <typename Real, class Op>
__device__ TileStage tile_stage(const Grid& g, const int* __restrict__ cs, int cell, TileShared<Real>& sh,
                                const Real* blo, const Real* bhi, const Real* wlo, const Real* whi,
                                Real R2, bool tile_act, Op& op)
{
    constexpr int CAP = Cap2<Real>::v;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int cx, cy, cz;
    cell_xyz(g, cell, cx, cy, cz);
    const int x0 = cx > 0 ? cx - 1 : 0, x1 = cx + 1 < g.dx ? cx + 1 : g.dx - 1;
    int count = 0, len = 0;
    bool over = false;
    for (int dz = -1; dz <= 1; dz++) {
        const int zz = cz + dz;
        if (zz < 0 || zz >= g.dz) continue;
        for (int dy = -1; dy <= 1; dy++) {
            const int yy = cy + dy;
            if (yy < 0 || yy >= g.dy) continue;
            const int row = zz * g.sheet + yy * g.dx;
            const int a = cs[row + x0], b = cs[row + x1 + 1];
            for (int base = a; base < b; base += TB) {
                const int j = base + tid;
                Cand<Real> c;
                bool keep = false;
                if (j < b) {
                    op.load_cand(j, c);
                    keep = box_d2(c.x, c.y, c.z, blo, bhi) < R2;
                }
                const uint64_t m = ballot(keep);
                if (lane == 0) sh.wcnt[wave] = __popcll(m);
                __syncthreads();
                int off = count, tot = 0;
#pragma unroll
                for (int w = 0; w < TW; w++) {
                    const int v = sh.wcnt[w];
                    off += w < wave ? v : 0;
                    tot += v;
                }
                if (keep) sh.S[off + lanes_below(m)] = c;
                count += tot;
                __syncthreads();
                if (count > CAP - TB) {
                    tile_batch(sh, count, wlo, whi, R2, tile_act, op);
                    count = 0;
                    over = true;
                }
            }
        }
    }
    if (count) len = tile_batch(sh, count, wlo, whi, R2, tile_act, op);
    return {count, len, over};
}

}  // namespace mps
