// mps_sort.hip -- cell hash, LDS-staged stable LSD radix sort, cell bounds and
// the reorder gather: the GPU form of MPS_GPU::divideCell + sort_* +
// sort_normal (MPS_CAL_expl_v1.0/MPS_GPU.cpp:248-373, MPS_CAL_/MPS_GPU.cpp:811-895).
//
// The reference hands out slots with a serial counting sort that is stable in
// the current order.  A stable LSD radix sort over the cell key with the
// current index as payload yields exactly that permutation.
#include <algorithm>

#include "mps_device.hpp"
#include "mps_kernels.hpp"

namespace mps {

// ---------------------------------------------------------------- hash ---
// c_a = integer((pos_a - origin_a) / h), truncation toward zero, taken in the
// state precision (MPS_CAL_expl_v1.0/MPS_GPU.cpp:265-268).  Out-of-grid
// particles become type 3 with zero velocity (MPS_GPU.cpp:270-280) and get
// key = ncells so that they sort after every cell.  With cnt, the keys are
// also counted per cell (cnt[ncells]: the excluded ones): keys arrive nearly
// sorted, so each run of equal keys in a wavefront adds its length once.
template <typename Real>
__global__ void __launch_bounds__(256) k_hash(Grid g, int n, const Real* __restrict__ x,
                                              const Real* __restrict__ y, const Real* __restrict__ z,
                                              int* __restrict__ type, Real* u, Real* v, Real* w,
                                              uint32_t* __restrict__ key, uint32_t* __restrict__ val,
                                              int* __restrict__ nexcl, uint32_t* __restrict__ cnt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    uint32_t kk = 0xffffffffu;
    if (valid) {
        val[i] = (uint32_t)i;
        kk = (uint32_t)g.ncells;
        if (type[i] != 3) {
            const Real left = (Real)g.left, back = (Real)g.back, bottom = (Real)g.bottom, h = (Real)g.h;
            Real fx = (x[i] - left) / h;
            Real fy = (y[i] - back) / h;
            Real fz = (z[i] - bottom) / h;
            Real tx = trunc(fx), ty = trunc(fy), tz = trunc(fz);
            bool in = tx >= (Real)0 && ty >= (Real)0 && tz >= (Real)0 && tx < (Real)g.dx &&
                      ty < (Real)g.dy && tz < (Real)g.dz;
            if (in) {
                kk = (uint32_t)((int)tz * g.sheet + (int)ty * g.dx + (int)tx);
            } else {
                type[i] = 3;
                u[i] = (Real)0; v[i] = (Real)0; w[i] = (Real)0;
                atomicAdd(nexcl, 1);
            }
        }
        key[i] = kk;
    }
    if (cnt) {
        const int lane = threadIdx.x & 63;
        const uint32_t prev = __shfl_up(kk, 1);
        const bool head = valid && (lane == 0 || prev != kk);
        const uint64_t bound = ballot(head || !valid);
        if (head) {
            const uint64_t up = lane == 63 ? 0ull : (bound >> (lane + 1));
            const int next = up ? lane + 1 + __builtin_ctzll(up) : 64;
            atomicAdd(&cnt[kk], (uint32_t)(next - lane));
        }
    }
}

// ------------------------------------------------- re-binning a running state ---
// Between two sorts of a running simulation a particle moves a fraction of a
// cell, so in the current (previously sorted) order the particles of new
// cell k lie in the 27 cells around k.  Walking those old cells in ascending
// index walks their slots in ascending order, so collecting the particles
// whose new key is k gives exactly the stable order within k.  One wavefront
// per new cell does that, writing into the cell's range of the scanned
// counts.  A cell that collects fewer particles than it counted (one came from
// further away, or left the grid) walks every slot instead: the result is the
// stable permutation in every case, only slower for such a cell.
__global__ void __launch_bounds__(64) k_gather_cells(Grid g, int n, const uint32_t* __restrict__ key,
                                                     const int* __restrict__ cs_old,
                                                     const uint32_t* __restrict__ cs_new,
                                                     uint32_t* __restrict__ perm)
{
    const int k = xcd_cell(g.ncells + 1);
    if (k > g.ncells) return;
    const int lane = threadIdx.x;
    const int lo = (int)cs_new[k], hi = k < g.ncells ? (int)cs_new[k + 1] : n;
    const int cnt = hi - lo;
    if (cnt <= 0) return;
    int found = 0;
    // slots [a, b) in ascending order, four 64-slot chunks loaded before the
    // first is matched (the loads are independent: one latency per four chunks)
    auto collect = [&](int a, int b) {
        for (int base = a; base < b; base += 4 * WAVE) {
            uint32_t kk[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int j = base + u * WAVE + lane;
                kk[u] = j < b ? key[j] : 0xffffffffu;
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int j = base + u * WAVE + lane;
                const bool hit = kk[u] == (uint32_t)k;
                const uint64_t m = ballot(hit);
                const int pos = found + lanes_below(m);
                if (hit && pos < cnt) perm[lo + pos] = (uint32_t)j;
                found += __popcll(m);
            }
        }
    };
    if (k < g.ncells) {
        int cx, cy, cz;
        cell_xyz(g, k, cx, cy, cz);
        const int x0 = cx > 0 ? cx - 1 : 0, x1 = cx + 1 < g.dx ? cx + 1 : g.dx - 1;
        // lane r < 9 reads the slot range of stencil row r (dz-major, as the walk)
        int ra = 0, rb = 0;
        if (lane < 9) {
            const int zz = cz + lane / 3 - 1, yy = cy + lane % 3 - 1;
            if (zz >= 0 && zz < g.dz && yy >= 0 && yy < g.dy) {
                const int row = zz * g.sheet + yy * g.dx;
                ra = cs_old[row + x0];
                rb = cs_old[row + x1 + 1];
            }
        }
        for (int r = 0; r < 9; r++) collect(__shfl(ra, r), __shfl(rb, r));
    } else {
        // the excluded tail: unchanged unless particles left the grid since
        const int a = cs_old[g.ncells];
        if (n - a == cnt) {
            for (int r = lane; r < cnt; r += WAVE) perm[lo + r] = (uint32_t)(a + r);
            found = cnt;
        }
    }
    if (found != cnt) {
        found = 0;
        collect(0, n);
    }
}

// ------------------------------------------------------------ radix sort ---
constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;

// Per-block digit histogram.  Keys arrive nearly sorted (one re-binning per
// step moves few particles), so equal digits come in runs: each run head adds
// the run length once instead of one LDS atomic per key.
__global__ void __launch_bounds__(RS_THREADS) k_radix_hist(const uint32_t* __restrict__ keys, int n,
                                                           int shift, uint32_t* __restrict__ hist)
{
    __shared__ uint32_t h[256];
    const int tid = threadIdx.x, lane = tid & 63;
    h[tid] = 0;
    __syncthreads();
    const int base = blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ITEMS; r++) {
        int i = base + r * RS_THREADS + tid;
        bool valid = i < n;
        uint32_t d = valid ? (keys[i] >> shift) & 255u : 0u;
        uint32_t dp = __shfl_up(d, 1);
        bool head = valid && (lane == 0 || dp != d);
        uint64_t bound = ballot(head || !valid);
        if (head) {
            uint64_t up = lane == 63 ? 0ull : (bound >> (lane + 1));
            int next = up ? lane + 1 + __builtin_ctzll(up) : 64;
            atomicAdd(&h[d], (uint32_t)(next - lane));
        }
    }
    __syncthreads();
    hist[(size_t)tid * gridDim.x + blockIdx.x] = h[tid];
}

// Exclusive scan of m counters in two launches: per-tile totals, then each
// tile rescanned with its offset.
constexpr int SC_THREADS = 256;
constexpr int SC_ITEMS = 16;
constexpr int SC_TILE = SC_THREADS * SC_ITEMS;

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* total)
{
    __shared__ uint32_t ws[SC_THREADS / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(inc, d);
        if (lane >= d) inc += t;
    }
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (int k = 0; k < SC_THREADS / 64; k++) {
        if (k < w) off += ws[k];
        tot += ws[k];
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

__global__ void __launch_bounds__(SC_THREADS) k_scan_tiles(const uint32_t* __restrict__ a, int m,
                                                           uint32_t* __restrict__ tiles)
{
    const int base = blockIdx.x * SC_TILE;
    uint32_t s = 0;
    for (int k = 0; k < SC_ITEMS; k++) {
        const int i = base + k * SC_THREADS + threadIdx.x;
        if (i < m) s += a[i];
    }
    uint32_t tot;
    block_excl_scan(s, &tot);
    if (threadIdx.x == 0) tiles[blockIdx.x] = tot;
}

// Each thread owns SC_ITEMS consecutive counters of the tile.  The tile's
// offset is the sum of the totals of the tiles before it, which every block
// forms itself (a few thousand tiles at most: no separate scan of the totals).
__global__ void __launch_bounds__(SC_THREADS) k_scan_apply(uint32_t* __restrict__ a, int m,
                                                           const uint32_t* __restrict__ tiles)
{
    uint32_t before = 0;
    for (int t = threadIdx.x; t < (int)blockIdx.x; t += SC_THREADS) before += tiles[t];
    uint32_t carry;
    block_excl_scan(before, &carry);
    const int base = blockIdx.x * SC_TILE + threadIdx.x * SC_ITEMS;
    uint32_t v[SC_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) {
        v[k] = base + k < m ? a[base + k] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t run = carry + block_excl_scan(s, &tot);
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) {
        if (base + k < m) a[base + k] = run;
        run += v[k];
    }
}

// Stable scatter.  Each round covers 256 consecutive keys; within a wave the
// lanes holding the same digit are found with 8 ballots, ranks follow lane
// order, and per-wave digit counts in LDS order the 4 waves.  Rounds advance a
// per-digit running offset, so the output order is the input order within
// each digit: the pass is stable.
__global__ void __launch_bounds__(RS_THREADS) k_radix_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    uint32_t* __restrict__ vout, int n, int shift, const uint32_t* __restrict__ offs)
{
    __shared__ uint32_t run[256];
    __shared__ uint32_t wc[4][256];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    run[tid] = offs[(size_t)tid * gridDim.x + blockIdx.x];
    wc[0][tid] = 0; wc[1][tid] = 0; wc[2][tid] = 0; wc[3][tid] = 0;
    __syncthreads();
    const int base = blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ITEMS; r++) {
        int i = base + r * RS_THREADS + tid;
        bool valid = i < n;
        uint32_t k = valid ? kin[i] : 0u;
        uint32_t v = valid ? vin[i] : 0u;
        uint32_t d = (k >> shift) & 255u;
        uint64_t m = ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            bool bit = (d >> b) & 1u;
            uint64_t bb = ballot(bit);
            m &= bit ? bb : ~bb;
        }
        int rank = lanes_below(m);
        if (valid && rank == 0) wc[wv][d] = (uint32_t)__popcll(m);
        __syncthreads();
        if (valid) {
            uint32_t pre = run[d];
            for (int q = 0; q < wv; q++) pre += wc[q][d];
            uint32_t dst = pre + (uint32_t)rank;
            kout[dst] = k;
            vout[dst] = v;
        }
        __syncthreads();
        run[tid] += wc[0][tid] + wc[1][tid] + wc[2][tid] + wc[3][tid];
        wc[0][tid] = 0; wc[1][tid] = 0; wc[2][tid] = 0; wc[3][tid] = 0;
        __syncthreads();
    }
}

// cell_start[c] = first sorted slot whose key >= c (the reference's exclusive
// scan of per-cell counts, MPS_GPU.cpp:293-298); cell_start[ncells] = number
// of particles still inside the grid.
__global__ void __launch_bounds__(256) k_cell_bounds(const uint32_t* __restrict__ keys, int n,
                                                     int ncells, int* __restrict__ cs)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    int k = i < n ? (int)keys[i] : ncells + 1;
    int kp = i == 0 ? -1 : (int)keys[i - 1];
    int hi = min(k, ncells);
    for (int c = kp + 1; c <= hi; c++) cs[c] = i;
}

// ------------------------------------------------------------- reorder ---
// Gather every per-particle array into sorted order (sort_i/d/d3,
// MPS_GPU.cpp:375-566) and record the inverse permutation for sort_normal.
template <typename Real>
__global__ void __launch_bounds__(256) k_reorder(int n, const uint32_t* __restrict__ perm,
                                                 StateView<Real> src, StateView<Real> dst,
                                                 int* __restrict__ inv, int* __restrict__ normal_tmp,
                                                 int* __restrict__ zero, int nzero)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    // clear the retired cell-start array: it counts the next re-binning's keys
    for (int c = k; c < nzero; c += gridDim.x * blockDim.x) zero[c] = 0;
    if (k >= n) return;
    int o = (int)perm[k];
    dst.x[k] = src.x[o]; dst.y[k] = src.y[o]; dst.z[k] = src.z[o];
    dst.u[k] = src.u[o]; dst.v[k] = src.v[o]; dst.w[k] = src.w[o];
    dst.p[k] = src.p[o]; dst.n[k] = src.n[o];
    dst.type[k] = src.type[o]; dst.id[k] = src.id[o];
    normal_tmp[k] = src.normal[o];
    inv[o] = k;
}

// i_normal = i_index[i_normal] (MPS_GPU.cpp:336-373)
__global__ void __launch_bounds__(256) k_normal_remap(int n, const int* __restrict__ normal_tmp,
                                                      const int* __restrict__ inv, int* __restrict__ normal)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int o = normal_tmp[k];
    normal[k] = o >= 0 ? inv[o] : -1;
}

// Excluded particles sit after every cell; the sweeps write cells only, so the
// tail of the ping-pong buffer they write into would hold another buffer's
// stale values.  After each sort both buffers get the sorted tail.
template <typename Real>
__global__ void __launch_bounds__(256) k_tail_copy(int n, const int* __restrict__ np_active,
                                                   StateView<Real> from, StateView<Real> to)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n || k < *np_active) return;
    to.x[k] = from.x[k]; to.y[k] = from.y[k]; to.z[k] = from.z[k];
    to.u[k] = from.u[k]; to.v[k] = from.v[k]; to.w[k] = from.w[k];
}

// ------------------------------------------------------------- drivers ---
static inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// radix histogram (256 digits x blocks) followed by the tile totals of its scan
int sort_scratch_words(int n)
{
    const int m = 256 * cdiv(n, RS_TILE);
    return m + cdiv(std::max(m, n + 1), SC_TILE) + 64;
}

static void scan_excl(uint32_t* a, int m, uint32_t* tiles, hipStream_t st)
{
    const int nt = cdiv(m, SC_TILE);
    hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(SC_THREADS), 0, st, a, m, tiles);
    hipLaunchKernelGGL(k_scan_apply, dim3(nt), dim3(SC_THREADS), 0, st, a, m, tiles);
}

template <typename Real>
void launch_hash(const Grid& g, int n, StateView<Real> s, uint32_t* key, uint32_t* val, int* nexcl,
                 uint32_t* cnt, hipStream_t st)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_hash<Real>, dim3(cdiv(n, 256)), dim3(256), 0, st, g, n, s.x, s.y, s.z,
                       s.type, s.u, s.v, s.w, key, val, nexcl, cnt);
}

int scan_tile_words(int m) { return cdiv(m, SC_TILE) + 64; }

void launch_gather_sort(const Grid& g, int n, const uint32_t* key, const int* cs_old, uint32_t* cs_new,
                        uint32_t* tiles, uint32_t* perm, hipStream_t st)
{
    if (n == 0) return;
    scan_excl(cs_new, g.ncells + 1, tiles, st);
    hipLaunchKernelGGL(k_gather_cells, dim3(xcd_grid(g.ncells + 1)), dim3(64), 0, st, g, n, key, cs_old,
                       (const uint32_t*)cs_new, perm);
}

void launch_radix_sort(int n, int key_bits, uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1,
                       uint32_t* hist, uint32_t** kres, uint32_t** vres, hipStream_t st)
{
    *kres = k0; *vres = v0;
    if (n == 0) return;
    const int nb = cdiv(n, RS_TILE);
    for (int shift = 0; shift < key_bits; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(RS_THREADS), 0, st, k0, n, shift, hist);
        scan_excl(hist, 256 * nb, hist + 256 * nb, st);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(RS_THREADS), 0, st, k0, v0, k1, v1, n,
                           shift, hist);
        uint32_t* t = k0; k0 = k1; k1 = t;
        t = v0; v0 = v1; v1 = t;
    }
    *kres = k0; *vres = v0;
}

void launch_scan_excl(uint32_t* a, int m, uint32_t* tiles, hipStream_t st) { scan_excl(a, m, tiles, st); }

void launch_cell_bounds(const uint32_t* keys, int n, int ncells, int* cs, hipStream_t st)
{
    hipLaunchKernelGGL(k_cell_bounds, dim3(cdiv(n + 1, 256)), dim3(256), 0, st, keys, n, ncells, cs);
}

template <typename Real>
void launch_reorder(int n, const uint32_t* perm, StateView<Real> src, StateView<Real> dst, int* inv,
                    int* normal_tmp, int* zero, int nzero, hipStream_t st)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_reorder<Real>, dim3(cdiv(n, 256)), dim3(256), 0, st, n, perm, src, dst, inv,
                       normal_tmp, zero, nzero);
    hipLaunchKernelGGL(k_normal_remap, dim3(cdiv(n, 256)), dim3(256), 0, st, n, normal_tmp, inv,
                       dst.normal);
}

template <typename Real>
void launch_tail_copy(int n, const int* np_active, StateView<Real> from, StateView<Real> to, hipStream_t st)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_tail_copy<Real>, dim3(cdiv(n, 256)), dim3(256), 0, st, n, np_active, from, to);
}

template void launch_tail_copy<float>(int, const int*, StateView<float>, StateView<float>, hipStream_t);
template void launch_tail_copy<double>(int, const int*, StateView<double>, StateView<double>, hipStream_t);
template void launch_hash<float>(const Grid&, int, StateView<float>, uint32_t*, uint32_t*, int*, uint32_t*,
                                  hipStream_t);
template void launch_hash<double>(const Grid&, int, StateView<double>, uint32_t*, uint32_t*, int*, uint32_t*,
                                   hipStream_t);
template void launch_reorder<float>(int, const uint32_t*, StateView<float>, StateView<float>, int*, int*, int*,
                                    int, hipStream_t);
template void launch_reorder<double>(int, const uint32_t*, StateView<double>, StateView<double>, int*, int*,
                                     int*, int, hipStream_t);

}  // namespace mps
