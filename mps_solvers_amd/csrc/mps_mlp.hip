// mps_mlp.hip -- learned initial guess for the pressure Poisson CG.
//
// No counterpart in the reference.  A per-particle MLP (16 -> 64 -> 64 -> 64
// -> 1, SiLU) maps local features of the assembled Poisson system to a
// correction of the previous-step pressure:  x0 = p_prev + P_s * y.
// Forward and backward passes run on the gfx950 matrix cores
// (v_mfma_f32_16x16x16_bf16: BF16 operands, FP32 accumulation); training is
// Adam on FP32 master weights.
//
// Matrix layout (16 x 16 tiles, lane l, q = l / 16, c = l % 16):
//   A operand: lane holds A[c][4q .. 4q+3];  B operand: lane holds B[4q .. 4q+3][c];
//   result:    lane holds D[4q .. 4q+3][c].
// Activations are kept transposed (feature x particle), so a layer's result
// tiles are directly the next layer's B operands:  Z_l^T = W_l^T H_{l-1}^T.
// Weight gradients need the particle dimension as K and go through a
// 16 x 16 transpose in LDS.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <random>
#include <vector>

#include "../../include/mps_amd.h"
#include "mps_ctx.hpp"
#include "mps_ppe.hpp"

namespace mps {

constexpr int NIN = 16, NH = 64;
constexpr int OW1 = 0, OB1 = 1024, OW2 = 1088, OB2 = 5184, OW3 = 5248, OB3 = 9344, OW4 = 9408, OB4 = 9472;
constexpr int NPARAM = 9473;
constexpr int MT = 256;   // threads per MLP block (4 waves, 16 samples per wave per pass)
constexpr int INFER_BLOCKS = 512;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short bf(float f)
{
    // round to nearest even
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}
__device__ __forceinline__ float bf_val(short h) { return __uint_as_float(((uint32_t)(unsigned short)h) << 16); }
__device__ __forceinline__ bf16x4 cvt4(const f32x4& v) { return bf16x4{bf(v[0]), bf(v[1]), bf(v[2]), bf(v[3])}; }
__device__ __forceinline__ f32x4 mfma(const bf16x4& a, const bf16x4& b, const f32x4& c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z)
{
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// Weights in LDS as BF16, [in][out] row-major.
struct MlpW {
    short w1[NIN * NH];
    short w2[NH * NH];
    short w3[NH * NH];
    float w4b[NH];   // bf16-rounded W4 as floats (vector-unit last layer)
    float b1[NH], b2[NH], b3[NH];
    float b4;
};

__device__ void load_weights(MlpW& W, const float* __restrict__ th)
{
    for (int k = threadIdx.x; k < NIN * NH; k += blockDim.x) W.w1[k] = bf(th[OW1 + k]);
    for (int k = threadIdx.x; k < NH * NH; k += blockDim.x) {
        W.w2[k] = bf(th[OW2 + k]);
        W.w3[k] = bf(th[OW3 + k]);
    }
    for (int k = threadIdx.x; k < NH; k += blockDim.x) {
        W.w4b[k] = bf_val(bf(th[OW4 + k]));
        W.b1[k] = th[OB1 + k]; W.b2[k] = th[OB2 + k]; W.b3[k] = th[OB3 + k];
    }
    if (threadIdx.x == 0) W.b4 = th[OB4];
    __syncthreads();
}

// A operand of W^T (out-tile o, in-block kb): lane holds W[kb*16 + 4q + j][o*16 + c]
__device__ __forceinline__ bf16x4 a_wt(const short* w, int ld, int o, int kb, int q, int c)
{
    const int base = (kb * 16 + 4 * q) * ld + o * 16 + c;
    return bf16x4{w[base], w[base + ld], w[base + 2 * ld], w[base + 3 * ld]};
}
// A operand of W (in-tile ib, out-block ob): lane holds W[ib*16 + c][ob*16 + 4q + j]
__device__ __forceinline__ bf16x4 a_w(const short* w, int ld, int ib, int ob, int q, int c)
{
    const short* p = w + (ib * 16 + c) * ld + ob * 16 + 4 * q;
    return bf16x4{p[0], p[1], p[2], p[3]};
}

struct Act {
    f32x4 z1[4], z2[4], z3[4];
    f32x4 h3[4];
    bf16x4 x;
    float y;
};

// Forward pass for 16 samples (particles c of this wave).  xrow: this lane's
// 4 features 4q .. 4q+3 of particle c (zeros when invalid).
__device__ __forceinline__ void mlp_forward(const MlpW& W, const f32x4& xin, Act& a)
{
    const int l = threadIdx.x & 63, q = l >> 4, c = l & 15;
    a.x = cvt4(xin);
    f32x4 h[4];
#pragma unroll
    for (int o = 0; o < 4; o++) {
        f32x4 z = mfma(a_wt(W.w1, NH, o, 0, q, c), a.x, f32x4{0, 0, 0, 0});
#pragma unroll
        for (int j = 0; j < 4; j++) z[j] += W.b1[o * 16 + 4 * q + j];
        a.z1[o] = z;
#pragma unroll
        for (int j = 0; j < 4; j++) h[o][j] = silu(z[j]);
    }
    bf16x4 hb[4];
#pragma unroll
    for (int k = 0; k < 4; k++) hb[k] = cvt4(h[k]);
#pragma unroll
    for (int o = 0; o < 4; o++) {
        f32x4 z = {0, 0, 0, 0};
#pragma unroll
        for (int kb = 0; kb < 4; kb++) z = mfma(a_wt(W.w2, NH, o, kb, q, c), hb[kb], z);
#pragma unroll
        for (int j = 0; j < 4; j++) z[j] += W.b2[o * 16 + 4 * q + j];
        a.z2[o] = z;
#pragma unroll
        for (int j = 0; j < 4; j++) h[o][j] = silu(z[j]);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) hb[k] = cvt4(h[k]);
#pragma unroll
    for (int o = 0; o < 4; o++) {
        f32x4 z = {0, 0, 0, 0};
#pragma unroll
        for (int kb = 0; kb < 4; kb++) z = mfma(a_wt(W.w3, NH, o, kb, q, c), hb[kb], z);
#pragma unroll
        for (int j = 0; j < 4; j++) z[j] += W.b3[o * 16 + 4 * q + j];
        a.z3[o] = z;
#pragma unroll
        for (int j = 0; j < 4; j++) a.h3[o][j] = silu(z[j]);
    }
    // last layer on the vector units: y_c = sum_k bf16(h3[k][c]) * bf16(w4[k]) + b4
    float s = 0.0f;
#pragma unroll
    for (int o = 0; o < 4; o++)
#pragma unroll
        for (int j = 0; j < 4; j++) s += bf_val(bf(a.h3[o][j])) * W.w4b[o * 16 + 4 * q + j];
    s += __shfl_xor(s, 16);
    s += __shfl_xor(s, 32);
    a.y = s + W.b4;
}

__device__ __forceinline__ f32x4 load_x(const float* __restrict__ X, int p, bool ok, int q)
{
    if (!ok) return f32x4{0, 0, 0, 0};
    const float* r = X + (size_t)p * NIN + 4 * q;
    return f32x4{r[0], r[1], r[2], r[3]};
}

// ------------------------------------------------------------- inference ---
// y for rows 0..n-1; with x_prev: out = x_prev + ps * y (0 on identity rows)
__global__ void __launch_bounds__(MT) k_mlp_infer(int n, const float* __restrict__ X, const float* __restrict__ th,
                                                  const unsigned char* __restrict__ ident,
                                                  const double* __restrict__ xprev, double ps,
                                                  float* __restrict__ yout, double* __restrict__ out)
{
    __shared__ MlpW W;
    load_weights(W, th);
    const int l = threadIdx.x & 63, q = l >> 4, c = l & 15;
    const int wave = blockIdx.x * (MT / 64) + (threadIdx.x >> 6), nwaves = gridDim.x * (MT / 64);
    for (int p0 = wave * 16; p0 < n; p0 += nwaves * 16) {
        const int p = p0 + c;
        Act a;
        mlp_forward(W, load_x(X, p, p < n, q), a);
        if (q == 0 && p < n) {
            if (yout) yout[p] = a.y;
            if (out) out[p] = ident && ident[p] ? 0.0 : xprev[p] + ps * (double)a.y;
        }
    }
}

// ------------------------------------------------------------- training ---
struct MlpTrainShared {
    MlpW W;
    float grad[NPARAM];
    short tr[MT / 64][2][16 * 16];   // per-wave transpose scratch (bf16)
    float loss;
};

// tile in result layout (lane holds T[4q+j][c]) -> A/B operand of its transpose
// (lane gets T[c][4q .. 4q+3])
__device__ __forceinline__ bf16x4 transpose_tile(short* s, const f32x4& v, int q, int c)
{
#pragma unroll
    for (int j = 0; j < 4; j++) s[(4 * q + j) * 16 + c] = bf(v[j]);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const short* r = s + c * 16 + 4 * q;
    bf16x4 out = bf16x4{r[0], r[1], r[2], r[3]};
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    return out;
}

__device__ __forceinline__ float sum16(float v)
{
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 4);
    v += __shfl_xor(v, 8);
    return v;
}

// gW[in][out] += sum_p H[in][p] dZ[out][p] for one in-tile / out-tile pair
__device__ __forceinline__ void acc_wgrad(float* g, int ld, int ib, int ob, const bf16x4& hT, const bf16x4& dzT,
                                          int q, int c)
{
    const f32x4 d = mfma(hT, dzT, f32x4{0, 0, 0, 0});
#pragma unroll
    for (int j = 0; j < 4; j++) atomicAdd(&g[(ib * 16 + 4 * q + j) * ld + ob * 16 + c], d[j]);
}

// Forward + backward of mean((y - t)^2) over the samples idx[0..B); each block
// writes its gradient and loss partial sums.
__global__ void __launch_bounds__(MT) k_mlp_train(int B, const int* __restrict__ idx, const float* __restrict__ X,
                                                  const float* __restrict__ T, const float* __restrict__ th,
                                                  float* __restrict__ gpart, float* __restrict__ lpart)
{
    __shared__ MlpTrainShared S;
    for (int k = threadIdx.x; k < NPARAM; k += MT) S.grad[k] = 0.0f;
    if (threadIdx.x == 0) S.loss = 0.0f;
    load_weights(S.W, th);
    const MlpW& W = S.W;
    const int l = threadIdx.x & 63, q = l >> 4, c = l & 15, wv = threadIdx.x >> 6;
    short* tA = S.tr[wv][0];
    short* tB = S.tr[wv][1];
    const float invB = 2.0f / (float)B;
    const int wave = blockIdx.x * (MT / 64) + wv, nwaves = gridDim.x * (MT / 64);
    float lsum = 0.0f;
    for (int s0 = wave * 16; s0 < B; s0 += nwaves * 16) {
        const int sidx = s0 + c;
        const bool ok = sidx < B;
        const int row = ok ? idx[sidx] : 0;
        Act a;
        const f32x4 xin = load_x(X, row, ok, q);
        mlp_forward(W, xin, a);
        const float t = ok ? T[row] : 0.0f;
        const float r = ok ? a.y - t : 0.0f;
        if (q == 0) lsum += r * r;
        const float dy = r * invB;
        // last layer
#pragma unroll
        for (int o = 0; o < 4; o++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const float gw = sum16(a.h3[o][j] * dy);
                if (c == 0) atomicAdd(&S.grad[OW4 + o * 16 + 4 * q + j], gw);
            }
        {
            const float gb = sum16(q == 0 ? dy : 0.0f);
            if (l == 0) atomicAdd(&S.grad[OB4], gb);
        }
        f32x4 dz[4];
#pragma unroll
        for (int o = 0; o < 4; o++)
#pragma unroll
            for (int j = 0; j < 4; j++) dz[o][j] = dy * W.w4b[o * 16 + 4 * q + j] * dsilu(a.z3[o][j]);
        // layers 3, 2: weight and bias gradients, then dH_{l-1} = W_l dZ_l
        for (int layer = 3; layer >= 2; layer--) {
            const short* w = layer == 3 ? W.w3 : W.w2;
            const int ow = layer == 3 ? OW3 : OW2, ob_ = layer == 3 ? OB3 : OB2;
            const f32x4* zin = layer == 3 ? a.z2 : a.z1;
            f32x4 hin[4];
#pragma unroll
            for (int k = 0; k < 4; k++)
#pragma unroll
                for (int j = 0; j < 4; j++) hin[k][j] = silu(zin[k][j]);
#pragma unroll
            for (int o = 0; o < 4; o++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const float gb = sum16(dz[o][j]);
                    if (c == 0) atomicAdd(&S.grad[ob_ + o * 16 + 4 * q + j], gb);
                }
            bf16x4 dzT[4], dzB[4];
#pragma unroll
            for (int o = 0; o < 4; o++) { dzT[o] = transpose_tile(tB, dz[o], q, c); dzB[o] = cvt4(dz[o]); }
#pragma unroll
            for (int ib = 0; ib < 4; ib++) {
                const bf16x4 hT = transpose_tile(tA, hin[ib], q, c);
#pragma unroll
                for (int o = 0; o < 4; o++) acc_wgrad(S.grad + ow, NH, ib, o, hT, dzT[o], q, c);
            }
            f32x4 dh[4];
#pragma unroll
            for (int ib = 0; ib < 4; ib++) {
                f32x4 d = {0, 0, 0, 0};
#pragma unroll
                for (int o = 0; o < 4; o++) d = mfma(a_w(w, NH, ib, o, q, c), dzB[o], d);
                dh[ib] = d;
            }
#pragma unroll
            for (int k = 0; k < 4; k++)
#pragma unroll
                for (int j = 0; j < 4; j++) dz[k][j] = dh[k][j] * dsilu(zin[k][j]);
        }
        // layer 1
#pragma unroll
        for (int o = 0; o < 4; o++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const float gb = sum16(dz[o][j]);
                if (c == 0) atomicAdd(&S.grad[OB1 + o * 16 + 4 * q + j], gb);
            }
        {
            f32x4 xf;
#pragma unroll
            for (int j = 0; j < 4; j++) xf[j] = bf_val(a.x[j]);
            const bf16x4 xT = transpose_tile(tA, xf, q, c);
#pragma unroll
            for (int o = 0; o < 4; o++) {
                const bf16x4 dzT = transpose_tile(tB, dz[o], q, c);
                acc_wgrad(S.grad + OW1, NH, 0, o, xT, dzT, q, c);
            }
        }
    }
    lsum = sum16(lsum);
    if (l == 0) atomicAdd(&S.loss, lsum);
    __syncthreads();
    for (int k = threadIdx.x; k < NPARAM; k += MT) gpart[(size_t)blockIdx.x * NPARAM + k] = S.grad[k];
    if (threadIdx.x == 0) lpart[blockIdx.x] = S.loss;
}

// Sum the block partials in a fixed order; optionally take one Adam step.
__global__ void __launch_bounds__(256) k_mlp_reduce_adam(int nb, int B, const float* __restrict__ gpart,
                                                         const float* __restrict__ lpart, float* __restrict__ grad,
                                                         float* __restrict__ th, float* __restrict__ m,
                                                         float* __restrict__ v, float lr, float b1c, float b2c,
                                                         int adam, double* __restrict__ loss)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) {
        double s = 0.0;
        for (int b = 0; b < nb; b++) s += lpart[b];
        *loss = s / (double)B;
    }
    if (k >= NPARAM) return;
    float g = 0.0f;
    for (int b = 0; b < nb; b++) g += gpart[(size_t)b * NPARAM + k];
    if (grad) grad[k] = g;
    if (!adam) return;
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    const float mk = beta1 * m[k] + (1.0f - beta1) * g;
    const float vk = beta2 * v[k] + (1.0f - beta2) * g * g;
    m[k] = mk;
    v[k] = vk;
    th[k] -= lr * (mk / b1c) / (sqrtf(vk / b2c) + eps);
}

// ------------------------------------------------------------- features ---
// Pass A (16 lanes per row): residual of the previous pressure r = b - A x,
// diagonal D, off-diagonal weight sum W, Jacobi correction c = r / D.
constexpr int FG = 16;
constexpr int FU = 4;   // entries per lane in flight in the feature passes

__global__ void __launch_bounds__(256) k_feat_a(int n, const int* __restrict__ rp, const int* __restrict__ col,
                                                const double* __restrict__ val, const double* __restrict__ b,
                                                const double* __restrict__ x, double* __restrict__ cvec,
                                                double* __restrict__ dvec, double* __restrict__ wvec,
                                                unsigned char* __restrict__ ident)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = t / FG, g = t % FG;
    if (row >= n) return;
    const int k0 = rp[row], k1 = rp[row + 1];
    double ax = 0.0, d = 0.0, w = 0.0;
    // FU entries per lane loaded in one predicated batch before their gathers
    // (as the CG's SpMV), added in ascending order
    for (int k = k0 + g; k < k1; k += FU * FG) {
        int jj[FU];
        double av[FU], xv[FU];
#pragma unroll
        for (int u = 0; u < FU; u++) {
            const int kk = k + u * FG;
            jj[u] = kk < k1 ? col[kk] : -1;
            av[u] = kk < k1 ? val[kk] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < FU; u++) xv[u] = jj[u] >= 0 ? x[jj[u]] : 0.0;
#pragma unroll
        for (int u = 0; u < FU; u++) {
            if (jj[u] < 0) continue;
            ax += av[u] * xv[u];
            if (jj[u] == row) d += av[u]; else w += av[u];
        }
    }
#pragma unroll
    for (int m = FG / 2; m >= 1; m >>= 1) {
        ax += __shfl_xor(ax, m);
        d += __shfl_xor(d, m);
        w += __shfl_xor(w, m);
    }
    if (g == 0) {
        const bool id = (k1 - k0) == 1 && d == 1.0;
        ident[row] = id ? 1 : 0;
        dvec[row] = d;
        wvec[row] = w;
        cvec[row] = id ? 0.0 : (b[row] - ax) / d;
    }
}

// Pass B: neighbour moments -> 16 features per row (zeros on identity rows).
template <typename Real>
__global__ void __launch_bounds__(256) k_feat_b(int n, const int* __restrict__ rp, const int* __restrict__ col,
                                                const double* __restrict__ val, const double* __restrict__ b,
                                                const double* __restrict__ x, const double* __restrict__ cvec,
                                                const double* __restrict__ dvec, const double* __restrict__ wvec,
                                                const unsigned char* __restrict__ ident, const double* __restrict__ nn,
                                                const Real* __restrict__ z, const int* __restrict__ type,
                                                const DevScalars* __restrict__ sc, double ps, double rhog, double dp,
                                                float* __restrict__ F)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = t / FG, g = t % FG;
    if (row >= n) return;
    const int k0 = rp[row], k1 = rp[row + 1];
    const double n0 = sc->n0;
    const double xi = x[row], ci = cvec[row], zi = (double)z[row];
    double sc_ = 0, sx = 0, sdev = 0, sxz = 0, szz = 0, scz = 0, sxx = 0;
    for (int k = k0 + g; k < k1; k += FU * FG) {
        int jj[FU];
        double av[FU];
#pragma unroll
        for (int u = 0; u < FU; u++) {
            const int kk = k + u * FG;
            const int j = kk < k1 ? col[kk] : row;
            jj[u] = j == row ? -1 : j;   // the diagonal and the slots past the row add nothing
            av[u] = kk < k1 ? val[kk] : 0.0;
        }
        double xj[FU], cj[FU], nj[FU], zj[FU];
#pragma unroll
        for (int u = 0; u < FU; u++) {
            const int j = jj[u] >= 0 ? jj[u] : row;
            xj[u] = x[j]; cj[u] = cvec[j]; nj[u] = nn[j]; zj[u] = (double)z[j];
        }
#pragma unroll
        for (int u = 0; u < FU; u++) {
            if (jj[u] < 0) continue;
            const double a = av[u];
            const double dz = zj[u] - zi, dx = xj[u] - xi, dc = cj[u] - ci;
            sc_ += a * cj[u];
            sx += a * xj[u];
            sdev += a * (nj[u] - n0) / n0;
            sxz += a * dx * dz;
            szz += a * dz * dz;
            scz += a * dc * dz;
            sxx += a * dx * dx;
        }
    }
#pragma unroll
    for (int m = FG / 2; m >= 1; m >>= 1) {
        sc_ += __shfl_xor(sc_, m); sx += __shfl_xor(sx, m); sdev += __shfl_xor(sdev, m);
        sxz += __shfl_xor(sxz, m); szz += __shfl_xor(szz, m); scz += __shfl_xor(scz, m);
        sxx += __shfl_xor(sxx, m);
    }
    if (g != 0) return;
    float* f = F + (size_t)row * NIN;
    if (ident[row]) {
        for (int k = 0; k < NIN; k++) f[k] = 0.0f;
        return;
    }
    const double d = dvec[row], w = wvec[row];
    const double iw = w > 0.0 ? 1.0 / w : 0.0, izz = szz > 0.0 ? 1.0 / szz : 0.0;
    f[0] = (float)(xi / ps);
    f[1] = (float)(ci / ps);
    f[2] = (float)(sc_ * iw / ps);
    f[3] = (float)((sx * iw - (w > 0.0 ? xi : 0.0)) / ps);
    f[4] = (float)(100.0 * (nn[row] - n0) / n0);
    f[5] = (float)(100.0 * sdev * iw);
    f[6] = (float)(b[row] / (d * ps));
    f[7] = (float)(w / fabs(d));
    f[8] = (float)(-sc_ / d / ps);
    f[9] = (float)(sxz * izz / rhog);
    f[10] = (float)(scz * izz * dp / ps);
    f[11] = (float)(fabs(d) / 60.0);
    f[12] = type[row] == 1 ? 1.0f : 0.0f;
    f[13] = (float)((k1 - k0) / 128.0);
    f[14] = (float)(sxx * iw / (ps * ps));
    f[15] = 1.0f;
}

// training targets: t = x_converged / P_s - f0  (f0 = x_prev / P_s); rows
// flagged by keep[] are copied (in row order) to the dataset.
__global__ void __launch_bounds__(256) k_keep(int n, const unsigned char* __restrict__ ident, uint32_t* __restrict__ keep)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keep[i] = ident[i] ? 0u : 1u;
    if (i == n) keep[n] = 0u;
}

__global__ void __launch_bounds__(256) k_append(int n, const uint32_t* __restrict__ pos, const unsigned char* __restrict__ ident,
                                                const float* __restrict__ F, const double* __restrict__ xconv, double ps,
                                                long base, long cap, float* __restrict__ DX, float* __restrict__ DT)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || ident[i]) return;
    const long d = base + (long)pos[i];
    if (d >= cap) return;
    for (int k = 0; k < NIN; k++) DX[d * NIN + k] = F[(size_t)i * NIN + k];
    DT[d] = (float)(xconv[i] / ps - (double)F[(size_t)i * NIN]);
}

}  // namespace mps

using namespace mps;

#define HIPOK(expr)                                                                           \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) {                                                               \
            fprintf(stderr, "mps_amd: %s failed: %s\n", #expr, hipGetErrorString(e_));      \
            return MPS_ERR_HIP;                                                               \
        }                                                                                     \
    } while (0)

struct MlpBuffers {
    float* th = nullptr;   // FP32 master parameters
    float *m = nullptr, *v = nullptr;
    float* gpart = nullptr;
    float* lpart = nullptr;
    float* grad = nullptr;
    double* loss = nullptr;
    int nb = 64;
    long adam_t = 0;
    bool ready = false;
    // features of the current Poisson system
    int nmax = 0, nrows = 0;
    float* F = nullptr;
    double *cvec = nullptr, *dvec = nullptr, *wvec = nullptr, *xprev = nullptr;
    unsigned char* ident = nullptr;
    uint32_t* keep = nullptr;
    bool have_features = false;
    // dataset
    long cap = 0, count = 0;
    float *DX = nullptr, *DT = nullptr;
    int* idx = nullptr;
    long idx_cap = 0;
    bool record = false;
};

static void mfree(void*& p)
{
    if (p) hipFree(p);
    p = nullptr;
}

void mps_mlp_free(mps_ctx* c)
{
    MlpBuffers* M = c->mlp;
    if (!M) return;
    void* ps[] = {M->th, M->m, M->v, M->gpart, M->lpart, M->grad, M->loss, M->F, M->cvec, M->dvec,
                  M->wvec, M->xprev, M->ident, M->keep, M->DX, M->DT, M->idx};
    for (void* p : ps) mfree(p);
    delete M;
    c->mlp = nullptr;
}

static int mlp_ensure(mps_ctx* c)
{
    if (!c->mlp) c->mlp = new MlpBuffers();
    MlpBuffers* M = c->mlp;
    if (M->th) return MPS_OK;
    HIPOK(hipMalloc(&M->th, 4 * NPARAM));
    HIPOK(hipMalloc(&M->m, 4 * NPARAM));
    HIPOK(hipMalloc(&M->v, 4 * NPARAM));
    HIPOK(hipMalloc(&M->grad, 4 * NPARAM));
    HIPOK(hipMalloc(&M->gpart, 4 * (size_t)NPARAM * M->nb));
    HIPOK(hipMalloc(&M->lpart, 4 * M->nb));
    HIPOK(hipMalloc(&M->loss, 8));
    HIPOK(hipMemset(M->th, 0, 4 * NPARAM));
    HIPOK(hipMemset(M->m, 0, 4 * NPARAM));
    HIPOK(hipMemset(M->v, 0, 4 * NPARAM));
    return MPS_OK;
}

static double mlp_ps(const mps_ctx* c)
{
    const double g = sqrt(c->P.g[0] * c->P.g[0] + c->P.g[1] * c->P.g[1] + c->P.g[2] * c->P.g[2]);
    return c->P.rho * (g > 0.0 ? g : 9.81) * 64.0 * c->P.dp;
}

extern "C" int mps_mlp_init(mps_ctx* c, unsigned long long seed)
{
    if (!c) return MPS_ERR_ARG;
    int rc = mlp_ensure(c);
    if (rc) return rc;
    std::mt19937_64 rng(seed);
    std::normal_distribution<float> nd(0.0f, 1.0f);
    std::vector<float> th(NPARAM, 0.0f);
    auto fill = [&](int off, int fan_in, int cnt) {
        const float s = 1.0f / sqrtf((float)fan_in);
        for (int k = 0; k < cnt; k++) th[off + k] = s * nd(rng);
    };
    fill(OW1, NIN, NIN * NH);
    fill(OW2, NH, NH * NH);
    fill(OW3, NH, NH * NH);
    fill(OW4, NH, NH);
    for (int k = 0; k < NPARAM; k++)
        if ((k >= OB1 && k < OW2) || (k >= OB2 && k < OW3) || (k >= OB3 && k < OW4) || k == OB4) th[k] = 0.0f;
    // the network starts as "no correction": the last layer is scaled down
    for (int k = 0; k < NH; k++) th[OW4 + k] *= 0.01f;
    HIPOK(hipMemcpy(c->mlp->th, th.data(), 4 * NPARAM, hipMemcpyHostToDevice));
    HIPOK(hipMemset(c->mlp->m, 0, 4 * NPARAM));
    HIPOK(hipMemset(c->mlp->v, 0, 4 * NPARAM));
    c->mlp->adam_t = 0;
    c->mlp->ready = true;
    return MPS_OK;
}

extern "C" int mps_mlp_set_params(mps_ctx* c, const float* theta)
{
    if (!c || !theta) return MPS_ERR_ARG;
    int rc = mlp_ensure(c);
    if (rc) return rc;
    HIPOK(hipMemcpy(c->mlp->th, theta, 4 * NPARAM, hipMemcpyHostToDevice));
    c->mlp->ready = true;
    return MPS_OK;
}

extern "C" int mps_mlp_get_params(mps_ctx* c, float* theta)
{
    if (!c || !theta || !c->mlp || !c->mlp->th) return MPS_ERR_STATE;
    HIPOK(hipStreamSynchronize(c->st));
    HIPOK(hipMemcpy(theta, c->mlp->th, 4 * NPARAM, hipMemcpyDeviceToHost));
    return MPS_OK;
}

extern "C" int mps_mlp_nparams(void) { return NPARAM; }

// Inference on caller features (n x 16, row-major): y (n).
extern "C" int mps_mlp_forward(mps_ctx* c, const float* X, int n, float* y)
{
    if (!c || !X || !y || n < 0) return MPS_ERR_ARG;
    if (!c->mlp || !c->mlp->ready) return MPS_ERR_STATE;
    if (n == 0) return MPS_OK;
    float *dX = nullptr, *dy = nullptr;
    HIPOK(hipMalloc(&dX, 4 * (size_t)n * NIN));
    HIPOK(hipMalloc(&dy, 4 * (size_t)n));
    HIPOK(hipMemcpy(dX, X, 4 * (size_t)n * NIN, hipMemcpyHostToDevice));
    const int grid = std::min(2048, (n + 63) / 64);
    hipLaunchKernelGGL(k_mlp_infer, dim3(grid), dim3(MT), 0, c->st, n, dX, c->mlp->th, nullptr, nullptr, 0.0,
                       dy, nullptr);
    HIPOK(hipGetLastError());
    HIPOK(hipStreamSynchronize(c->st));
    HIPOK(hipMemcpy(y, dy, 4 * (size_t)n, hipMemcpyDeviceToHost));
    hipFree(dX);
    hipFree(dy);
    return MPS_OK;
}

static int run_train(mps_ctx* c, int B, const int* didx, const float* dX, const float* dT, int adam, float lr,
                     double* loss)
{
    MlpBuffers* M = c->mlp;
    const int nb = std::min(M->nb, std::max(1, (B + 63) / 64));
    hipLaunchKernelGGL(k_mlp_train, dim3(nb), dim3(MT), 0, c->st, B, didx, dX, dT, M->th, M->gpart, M->lpart);
    float b1c = 1.0f, b2c = 1.0f;
    if (adam) {
        M->adam_t++;
        b1c = 1.0f - powf(0.9f, (float)M->adam_t);
        b2c = 1.0f - powf(0.999f, (float)M->adam_t);
    }
    hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, c->st, nb, B, M->gpart,
                       M->lpart, M->grad, M->th, M->m, M->v, lr, b1c, b2c, adam, M->loss);
    HIPOK(hipGetLastError());
    if (loss) {
        HIPOK(hipMemcpyAsync(loss, M->loss, 8, hipMemcpyDeviceToHost, c->st));
        HIPOK(hipStreamSynchronize(c->st));
    }
    return MPS_OK;
}

// Gradient of mean((y - t)^2) over caller samples (no parameter update).
extern "C" int mps_mlp_grad(mps_ctx* c, const float* X, const float* t, int n, float* grad, double* loss)
{
    if (!c || !X || !t || !grad || n <= 0) return MPS_ERR_ARG;
    if (!c->mlp || !c->mlp->ready) return MPS_ERR_STATE;
    float *dX = nullptr, *dT = nullptr;
    int* di = nullptr;
    HIPOK(hipMalloc(&dX, 4 * (size_t)n * NIN));
    HIPOK(hipMalloc(&dT, 4 * (size_t)n));
    HIPOK(hipMalloc(&di, 4 * (size_t)n));
    std::vector<int> id((size_t)n);
    for (int i = 0; i < n; i++) id[i] = i;
    HIPOK(hipMemcpy(dX, X, 4 * (size_t)n * NIN, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(dT, t, 4 * (size_t)n, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(di, id.data(), 4 * (size_t)n, hipMemcpyHostToDevice));
    double l = 0.0;
    int rc = run_train(c, n, di, dX, dT, 0, 0.0f, &l);
    if (!rc) {
        HIPOK(hipMemcpy(grad, c->mlp->grad, 4 * NPARAM, hipMemcpyDeviceToHost));
        if (loss) *loss = l;
    }
    hipFree(dX);
    hipFree(dT);
    hipFree(di);
    return rc;
}

double* mps_ppe_guess_buffer(mps_ctx* c, int n);

struct PpeView {
    int n;
    const int* rp;
    const int* col;
    const double* v64;
    const double* b;
    const double* x;
    const unsigned char* bc;
};
int mps_ppe_view(mps_ctx* c, PpeView* v);   // mps_ppe_capi.hip

static int feat_ensure(mps_ctx* c, int n)
{
    MlpBuffers* M = c->mlp;
    if (M->nmax >= n && M->F) return MPS_OK;
    void* olds[] = {M->F, M->cvec, M->dvec, M->wvec, M->xprev, M->ident, M->keep};
    for (void* p : olds) if (p) hipFree(p);
    const size_t N = (size_t)std::max(n, 1);
    HIPOK(hipMalloc(&M->F, 4 * N * NIN));
    HIPOK(hipMalloc(&M->cvec, 8 * N));
    HIPOK(hipMalloc(&M->dvec, 8 * N));
    HIPOK(hipMalloc(&M->wvec, 8 * N));
    HIPOK(hipMalloc(&M->xprev, 8 * N));
    HIPOK(hipMalloc(&M->ident, N));
    HIPOK(hipMalloc(&M->keep, 4 * (N + 1)));
    M->nmax = n;
    return MPS_OK;
}

__global__ void k_xprev_f(int n, const float* __restrict__ p, double* __restrict__ x)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = (double)p[i];
}
__global__ void k_xprev_d(int n, const double* __restrict__ p, double* __restrict__ x)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = p[i];
}

// Features of the assembled system, with x_prev = the state pressure.
static int compute_features(mps_ctx* c)
{
    int rc = mlp_ensure(c);
    if (rc) return rc;
    PpeView v;
    if ((rc = mps_ppe_view(c, &v))) return rc;
    if ((rc = feat_ensure(c, v.n))) return rc;
    MlpBuffers* M = c->mlp;
    const int n = v.n;
    const int g1 = (n + 255) / 256;
    if (c->prec) hipLaunchKernelGGL(k_xprev_d, dim3(g1), dim3(256), 0, c->st, n, (const double*)c->f[c->aux][6], M->xprev);
    else hipLaunchKernelGGL(k_xprev_f, dim3(g1), dim3(256), 0, c->st, n, (const float*)c->f[c->aux][6], M->xprev);
    const int gr = (int)(((long)n * FG + 255) / 256);
    hipLaunchKernelGGL(k_feat_a, dim3(gr), dim3(256), 0, c->st, n, v.rp, v.col, v.v64, v.b, M->xprev, M->cvec,
                       M->dvec, M->wvec, M->ident);
    const double ps = mlp_ps(c);
    const double g = sqrt(c->P.g[0] * c->P.g[0] + c->P.g[1] * c->P.g[1] + c->P.g[2] * c->P.g[2]);
    const double rhog = c->P.rho * (g > 0.0 ? g : 9.81);
    if (c->prec)
        hipLaunchKernelGGL(k_feat_b<double>, dim3(gr), dim3(256), 0, c->st, n, v.rp, v.col, v.v64, v.b, M->xprev,
                           M->cvec, M->dvec, M->wvec, M->ident, c->n[c->aux], (const double*)c->f[c->kin][2],
                           c->type[c->aux], c->sc, ps, rhog, c->P.dp, M->F);
    else
        hipLaunchKernelGGL(k_feat_b<float>, dim3(gr), dim3(256), 0, c->st, n, v.rp, v.col, v.v64, v.b, M->xprev,
                           M->cvec, M->dvec, M->wvec, M->ident, c->n[c->aux], (const float*)c->f[c->kin][2],
                           c->type[c->aux], c->sc, ps, rhog, c->P.dp, M->F);
    HIPOK(hipGetLastError());
    M->nrows = n;
    M->have_features = true;
    return MPS_OK;
}

int mps_mlp_prepare(mps_ctx* c) { return compute_features(c); }

extern "C" int mps_mlp_features(mps_ctx* c, float* F, int* nrows)
{
    if (!c) return MPS_ERR_ARG;
    int rc = compute_features(c);
    if (rc) return rc;
    HIPOK(hipStreamSynchronize(c->st));
    if (nrows) *nrows = c->mlp->nrows;
    if (F) HIPOK(hipMemcpy(F, c->mlp->F, 4 * (size_t)c->mlp->nrows * NIN, hipMemcpyDeviceToHost));
    return MPS_OK;
}

// learned guess into the solver's guess buffer (x0_kind 3)
int mps_mlp_guess(mps_ctx* c)
{
    if (!c->mlp || !c->mlp->ready) return MPS_ERR_STATE;
    int rc = compute_features(c);
    if (rc) return rc;
    MlpBuffers* M = c->mlp;
    double* out = mps_ppe_guess_buffer(c, M->nrows);
    if (!out) return MPS_ERR_HIP;
    // a few blocks per CU: each block loads the weights into LDS once and then
    // streams its rows (the load is the fixed cost of a block)
    const int grid = std::min(INFER_BLOCKS, (M->nrows + 63) / 64);
    hipLaunchKernelGGL(k_mlp_infer, dim3(std::max(grid, 1)), dim3(MT), 0, c->st, M->nrows, M->F, M->th, M->ident,
                       M->xprev, mlp_ps(c), nullptr, out);
    HIPOK(hipGetLastError());
    return MPS_OK;
}

extern "C" int mps_mlp_infer(mps_ctx* c)
{
    if (!c) return MPS_ERR_ARG;
    int rc = mps_mlp_guess(c);
    if (rc) return rc;
    HIPOK(hipStreamSynchronize(c->st));
    return MPS_OK;
}

static int dataset_reserve(mps_ctx* c, long need);

// Append (features, converged correction) of the last solve to the dataset.
int mps_mlp_record(mps_ctx* c)
{
    MlpBuffers* M = c->mlp;
    if (!M || !M->have_features) return MPS_ERR_STATE;
    PpeView v;
    int rc = mps_ppe_view(c, &v);
    if (rc) return rc;
    if (v.n != M->nrows) return MPS_ERR_STATE;
    const int n = v.n;
    hipLaunchKernelGGL(k_keep, dim3((n + 256) / 256), dim3(256), 0, c->st, n, M->ident, M->keep);
    launch_scan_excl(M->keep, n + 1, (uint32_t*)M->cvec, c->st);
    uint32_t add = 0;
    HIPOK(hipMemcpyAsync(&add, M->keep + n, 4, hipMemcpyDeviceToHost, c->st));
    HIPOK(hipStreamSynchronize(c->st));
    if ((rc = dataset_reserve(c, M->count + (long)add))) return rc;
    hipLaunchKernelGGL(k_append, dim3((n + 255) / 256), dim3(256), 0, c->st, n, M->keep, M->ident, M->F, v.x,
                       mlp_ps(c), M->count, M->cap, M->DX, M->DT);
    HIPOK(hipGetLastError());
    M->count += add;
    return MPS_OK;
}

extern "C" int mps_mlp_dataset(mps_ctx* c, long* count, float* X, float* T)
{
    if (!c || !count) return MPS_ERR_ARG;
    *count = c->mlp ? c->mlp->count : 0;
    if (!c->mlp || !c->mlp->count) return MPS_OK;
    HIPOK(hipStreamSynchronize(c->st));
    if (X) HIPOK(hipMemcpy(X, c->mlp->DX, 4 * (size_t)c->mlp->count * NIN, hipMemcpyDeviceToHost));
    if (T) HIPOK(hipMemcpy(T, c->mlp->DT, 4 * (size_t)c->mlp->count, hipMemcpyDeviceToHost));
    return MPS_OK;
}

static int dataset_reserve(mps_ctx* c, long need)
{
    MlpBuffers* M = c->mlp;
    if (need <= M->cap) return MPS_OK;
    const long ncap = std::max((long)(need * 1.5), 1L << 16);
    float *nx = nullptr, *nt = nullptr;
    HIPOK(hipMalloc(&nx, 4 * (size_t)ncap * NIN));
    HIPOK(hipMalloc(&nt, 4 * (size_t)ncap));
    if (M->count) {
        HIPOK(hipMemcpy(nx, M->DX, 4 * (size_t)M->count * NIN, hipMemcpyDeviceToDevice));
        HIPOK(hipMemcpy(nt, M->DT, 4 * (size_t)M->count, hipMemcpyDeviceToDevice));
    }
    mfree((void*&)M->DX);
    mfree((void*&)M->DT);
    M->DX = nx;
    M->DT = nt;
    M->cap = ncap;
    return MPS_OK;
}

// Append caller samples (n x 16 features, n targets) to the training set.
extern "C" int mps_mlp_dataset_append(mps_ctx* c, const float* X, const float* T, long n)
{
    if (!c || !X || !T || n < 0) return MPS_ERR_ARG;
    int rc = mlp_ensure(c);
    if (rc) return rc;
    MlpBuffers* M = c->mlp;
    if ((rc = dataset_reserve(c, M->count + n))) return rc;
    HIPOK(hipStreamSynchronize(c->st));
    HIPOK(hipMemcpy(M->DX + M->count * NIN, X, 4 * (size_t)n * NIN, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(M->DT + M->count, T, 4 * (size_t)n, hipMemcpyHostToDevice));
    M->count += n;
    return MPS_OK;
}

extern "C" int mps_mlp_dataset_clear(mps_ctx* c)
{
    if (!c) return MPS_ERR_ARG;
    if (c->mlp) c->mlp->count = 0;
    return MPS_OK;
}

extern "C" int mps_mlp_set_record(mps_ctx* c, int on)
{
    if (!c) return MPS_ERR_ARG;
    int rc = mlp_ensure(c);
    if (rc) return rc;
    c->mlp->record = on != 0;
    return MPS_OK;
}

int mps_mlp_recording(mps_ctx* c) { return c->mlp && c->mlp->record; }

// iters Adam steps on minibatches of `batch` samples drawn without
// replacement (reshuffled every pass over the data).
extern "C" int mps_mlp_train(mps_ctx* c, int iters, int batch, double lr, unsigned long long seed,
                             double* loss_first, double* loss_last)
{
    if (!c || iters < 0 || batch <= 0) return MPS_ERR_ARG;
    if (!c->mlp || !c->mlp->ready) return MPS_ERR_STATE;
    MlpBuffers* M = c->mlp;
    if (M->count < 16) return MPS_ERR_STATE;
    const long N = M->count;
    batch = (int)std::min<long>(batch, N);
    if (M->idx_cap < N) {
        mfree((void*&)M->idx);
        HIPOK(hipMalloc(&M->idx, 4 * (size_t)N));
        M->idx_cap = N;
    }
    std::mt19937_64 rng(seed);
    std::vector<int> perm((size_t)N);
    for (long i = 0; i < N; i++) perm[i] = (int)i;
    long pos = N;
    for (int it = 0; it < iters; it++) {
        if (pos + batch > N) {
            std::shuffle(perm.begin(), perm.end(), rng);
            HIPOK(hipMemcpyAsync(M->idx, perm.data(), 4 * (size_t)N, hipMemcpyHostToDevice, c->st));
            HIPOK(hipStreamSynchronize(c->st));
            pos = 0;
        }
        double l = 0.0;
        const bool want = (it == 0 && loss_first) || (it == iters - 1 && loss_last);
        int rc = run_train(c, batch, M->idx + pos, M->DX, M->DT, 1, (float)lr, want ? &l : nullptr);
        if (rc) return rc;
        if (it == 0 && loss_first) *loss_first = l;
        if (it == iters - 1 && loss_last) *loss_last = l;
        pos += batch;
    }
    HIPOK(hipStreamSynchronize(c->st));
    return MPS_OK;
}
