// mps_ppe_capi.hip -- C ABI of the pressure Poisson solve and the
// semi-implicit step (MPS_CPU_EXPL::step, MPS_CAL_v4.0/MPS_CPU_EXPL.cpp:79-135).
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/mps_amd.h"
#include "mps_ctx.hpp"
#include "mps_ppe.hpp"

using namespace mps;

#define HIPOK(expr)                                                                           \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) {                                                               \
            fprintf(stderr, "mps_amd: %s failed: %s\n", #expr, hipGetErrorString(e_));      \
            return MPS_ERR_HIP;                                                               \
        }                                                                                     \
    } while (0)

struct PpeBuffers {
    int nmax = 0;        // rows allocated
    int n = 0;           // rows of the assembled system
    long cap = 0, nnz = 0;
    int* rp = nullptr;
    int* col = nullptr;
    float* v32 = nullptr;
    double* v64 = nullptr;
    double *b = nullptr, *x = nullptr, *x0_ext = nullptr;
    unsigned char* bc = nullptr;
    void *e = nullptr, *r = nullptr, *p = nullptr, *q = nullptr;
    double *partial = nullptr, *partial2 = nullptr;
    PpeScalars* ps = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;   // solve timing of mps_step_semi
    bool assembled = false;
};

static void ffree(void*& p)
{
    if (p) hipFree(p);
    p = nullptr;
}

void mps_ppe_free(mps_ctx* c)
{
    PpeBuffers* B = c->ppe;
    if (!B) return;
    void* ptrs[] = {B->rp, B->col, B->v32, B->v64, B->b, B->x, B->x0_ext, B->bc, B->e, B->r, B->p,
                    B->q, B->partial, B->partial2, B->ps};
    for (void* p : ptrs) ffree(p);
    if (B->t0) hipEventDestroy(B->t0);
    if (B->t1) hipEventDestroy(B->t1);
    delete B;
    c->ppe = nullptr;
}

static int ensure_rows(mps_ctx* c, int n)
{
    if (!c->ppe) c->ppe = new PpeBuffers();
    PpeBuffers* B = c->ppe;
    if (B->nmax >= n && B->ps) return MPS_OK;
    void* olds[] = {B->rp, B->b, B->x, B->x0_ext, B->bc, B->e, B->r, B->p, B->q, B->partial,
                    B->partial2};
    for (void* p : olds) if (p) hipFree(p);
    const size_t N = (size_t)std::max(n, 1);
    const size_t nb = (size_t)ppe_partial_blocks((int)N);
    HIPOK(hipMalloc(&B->rp, 4 * (N + 1)));
    HIPOK(hipMalloc(&B->b, 8 * N));
    HIPOK(hipMalloc(&B->x, 8 * N));
    HIPOK(hipMalloc(&B->x0_ext, 8 * N));
    HIPOK(hipMalloc(&B->bc, N));
    HIPOK(hipMalloc(&B->e, 8 * N));
    HIPOK(hipMalloc(&B->r, 8 * N));
    HIPOK(hipMalloc(&B->p, 8 * N));
    HIPOK(hipMalloc(&B->q, 8 * N));
    HIPOK(hipMalloc(&B->partial, 8 * nb));
    HIPOK(hipMalloc(&B->partial2, 8 * nb));
    if (!B->ps) {
        HIPOK(hipMalloc(&B->ps, sizeof(PpeScalars)));
        HIPOK(hipMemset(B->ps, 0, sizeof(PpeScalars)));
        HIPOK(hipEventCreate(&B->t0));
        HIPOK(hipEventCreate(&B->t1));
    }
    HIPOK(hipMemset(B->x0_ext, 0, 8 * N));
    B->nmax = n;
    return MPS_OK;
}

static int ensure_cap(mps_ctx* c, long nnz)
{
    PpeBuffers* B = c->ppe;
    if (B->cap >= nnz && B->col) return MPS_OK;
    ffree((void*&)B->col);
    ffree((void*&)B->v32);
    ffree((void*&)B->v64);
    const long cap = (long)(nnz * 1.15) + 1024;
    HIPOK(hipMalloc(&B->col, 4 * (size_t)cap));
    HIPOK(hipMalloc(&B->v32, 4 * (size_t)cap));
    HIPOK(hipMalloc(&B->v64, 8 * (size_t)cap));
    B->cap = cap;
    return MPS_OK;
}

template <typename Real>
static int assemble_t(mps_ctx* c, int n, double alpha, long* nnz)
{
    PpeBuffers* B = c->ppe;
    StateView<Real> s = c->view<Real>(c->kin, c->aux);
    const Thresh<Real>& th = *reinterpret_cast<const Thresh<Real>*>(c->prec ? (void*)&c->th.d : (void*)&c->th.f);
    const double d = (double)c->P.dim;
    // b_i = lambda rho / (2 d dt^2) (n0 - n_i); A_ii -= rho lambda alpha n0 / (2 d dt^2)
    const double btmp_over = c->P.rho / (2.0 * d);
    const double aii_over = c->P.rho * alpha / (2.0 * d);
    launch_ppe_rows<Real>(n, s, c->sc, c->P.beta, btmp_over, 1, nullptr, B->bc, B->b, B->x, B->partial,
                          B->ps, c->st);
    launch_ppe_assemble<Real>(c->g, c->cs, s, th, B->bc, c->sc, aii_over, 0, B->rp, nullptr, nullptr,
                              nullptr, c->st);
    HIPOK(hipMemsetAsync(B->rp + n, 0, 4, c->st));
    launch_scan_excl((uint32_t*)B->rp, n + 1, (uint32_t*)B->partial, c->st);
    int tot = 0;
    HIPOK(hipMemcpyAsync(&tot, B->rp + n, 4, hipMemcpyDeviceToHost, c->st));
    HIPOK(hipStreamSynchronize(c->st));
    int rc = ensure_cap(c, tot);
    if (rc) return rc;
    launch_ppe_assemble<Real>(c->g, c->cs, s, th, B->bc, c->sc, aii_over, 1, B->rp, B->col, B->v32,
                              B->v64, c->st);
    HIPOK(hipGetLastError());
    B->n = n;
    B->nnz = tot;
    B->assembled = true;
    if (nnz) *nnz = tot;
    return MPS_OK;
}

extern "C" int mps_ppe_assemble(mps_ctx* c, double alpha, long* nnz)
{
    if (!c) return MPS_ERR_ARG;
    int na = 0;
    int rc = mps_get_np(c, nullptr, &na);
    if (rc) return rc;
    if ((rc = ensure_rows(c, na))) return rc;
    return c->prec ? assemble_t<double>(c, na, alpha, nnz) : assemble_t<float>(c, na, alpha, nnz);
}

extern "C" int mps_ppe_get(mps_ctx* c, int* row_ptr, int* col, double* val, double* b, double* x,
                           unsigned char* bc)
{
    if (!c || !c->ppe || !c->ppe->assembled) return MPS_ERR_STATE;
    PpeBuffers* B = c->ppe;
    HIPOK(hipStreamSynchronize(c->st));
    const size_t n = (size_t)B->n;
    if (row_ptr) HIPOK(hipMemcpy(row_ptr, B->rp, 4 * (n + 1), hipMemcpyDeviceToHost));
    if (col) HIPOK(hipMemcpy(col, B->col, 4 * (size_t)B->nnz, hipMemcpyDeviceToHost));
    if (val) HIPOK(hipMemcpy(val, B->v64, 8 * (size_t)B->nnz, hipMemcpyDeviceToHost));
    if (b) HIPOK(hipMemcpy(b, B->b, 8 * n, hipMemcpyDeviceToHost));
    if (x) HIPOK(hipMemcpy(x, B->x, 8 * n, hipMemcpyDeviceToHost));
    if (bc) HIPOK(hipMemcpy(bc, B->bc, n, hipMemcpyDeviceToHost));
    return MPS_OK;
}

extern "C" int mps_ppe_info(mps_ctx* c, int* nrows, long* nnz)
{
    if (!c || !c->ppe || !c->ppe->assembled) return MPS_ERR_STATE;
    if (nrows) *nrows = c->ppe->n;
    if (nnz) *nnz = c->ppe->nnz;
    return MPS_OK;
}

__global__ void k_ppe_x0(int n, int kind, const float* __restrict__ pf, const double* __restrict__ pd,
                         const double* __restrict__ ext, const unsigned char* __restrict__ bc,
                         double* __restrict__ x)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = 0.0;
    if (kind == 1) v = pf ? (double)pf[i] : pd[i];
    else if (kind >= 2) v = ext[i];
    x[i] = bc[i] ? 0.0 : v;
}

template <typename V>
static int solve_t(mps_ctx* c, double tol, double tol_in, int maxit, int* iters, double* relres)
{
    PpeBuffers* B = c->ppe;
    const int n = B->n;
    V *e = (V*)B->e, *r = (V*)B->r, *p = (V*)B->p, *q = (V*)B->q;
    PpeScalars hs;
    int total = 0;
    double rel = INFINITY;
    const int batch = 16;
    for (int outer = 0; outer < 40; outer++) {
        // floor: the inner solve need not go below half the target on |b|
        HIPOK(hipMemcpy(&hs, B->ps, sizeof hs, hipMemcpyDeviceToHost));
        const double floor2 = 0.25 * tol * tol * hs.bb;
        launch_residual<V>(n, B->rp, B->col, B->v64, B->b, B->x, e, r, p, B->partial, B->partial2, B->ps,
                           tol_in * tol_in, floor2, c->st);
        HIPOK(hipGetLastError());
        HIPOK(hipStreamSynchronize(c->st));
        HIPOK(hipMemcpy(&hs, B->ps, sizeof hs, hipMemcpyDeviceToHost));
        rel = hs.bb > 0.0 ? sqrt(hs.rr64 / hs.bb) : sqrt(hs.rr64);
        if (rel <= tol || hs.bb == 0.0) break;
        if (total >= maxit) break;
        while (!hs.done && total + hs.iters < maxit) {
            for (int k = 0; k < batch; k++) {
                if (sizeof(V) == 4)
                    launch_cg_iter<float, float>(n, B->rp, B->col, B->v32, (float*)e, (float*)r, (float*)p,
                                                 (float*)q, B->partial, B->partial2, B->ps, c->st);
                else
                    launch_cg_iter<double, double>(n, B->rp, B->col, B->v64, (double*)e, (double*)r,
                                                   (double*)p, (double*)q, B->partial, B->partial2, B->ps,
                                                   c->st);
            }
            HIPOK(hipGetLastError());
            HIPOK(hipStreamSynchronize(c->st));
            HIPOK(hipMemcpy(&hs, B->ps, sizeof hs, hipMemcpyDeviceToHost));
        }
        total += hs.iters;
        launch_ir_update<V>(n, B->x, e, c->st);
        if (hs.done == 2) break;   // breakdown: p.Ap <= 0
    }
    HIPOK(hipStreamSynchronize(c->st));
    if (iters) *iters = total;
    if (relres) *relres = rel;
    return rel <= tol ? MPS_OK : MPS_ERR_STATE;
}

// x0_kind: 0 zero, 1 current pressure (previous step), 2 host array x0, 3 device guess
// left by mps_mlp_infer.  mixed: 1 FP32 inner CG + FP64 refinement; 0 FP64 CG.
extern "C" int mps_ppe_solve(mps_ctx* c, int mixed, int x0_kind, const double* x0, double tol,
                             int maxit, int* iters, double* relres)
{
    if (!c || !c->ppe || !c->ppe->assembled || !(tol > 0.0) || x0_kind < 0 || x0_kind > 3) return MPS_ERR_ARG;
    PpeBuffers* B = c->ppe;
    const int n = B->n;
    if (x0_kind == 2) {
        if (!x0) return MPS_ERR_ARG;
        HIPOK(hipMemcpyAsync(B->x0_ext, x0, 8 * (size_t)n, hipMemcpyHostToDevice, c->st));
    }
    const float* pf = c->prec ? nullptr : (const float*)c->f[c->aux][6];
    const double* pd = c->prec ? (const double*)c->f[c->aux][6] : nullptr;
    hipLaunchKernelGGL(k_ppe_x0, dim3((n + 255) / 256), dim3(256), 0, c->st, n, x0_kind, pf, pd,
                       B->x0_ext, B->bc, B->x);
    HIPOK(hipGetLastError());
    return mixed ? solve_t<float>(c, tol, 1e-4, maxit, iters, relres)
                 : solve_t<double>(c, tol, 0.0, maxit, iters, relres);
}

extern "C" int mps_ppe_get_x(mps_ctx* c, double* x)
{
    return mps_ppe_get(c, nullptr, nullptr, nullptr, nullptr, x, nullptr);
}

struct PpeView {
    int n;
    const int* rp;
    const int* col;
    const double* v64;
    const double* b;
    const double* x;
    const unsigned char* bc;
};

int mps_ppe_view(mps_ctx* c, PpeView* v)
{
    if (!c->ppe || !c->ppe->assembled) return MPS_ERR_STATE;
    PpeBuffers* B = c->ppe;
    v->n = B->n; v->rp = B->rp; v->col = B->col; v->v64 = B->v64; v->b = B->b; v->x = B->x; v->bc = B->bc;
    return MPS_OK;
}

double* mps_ppe_guess_buffer(mps_ctx* c, int n)
{
    if (ensure_rows(c, n)) return nullptr;
    return c->ppe->x0_ext;
}

// pressCorr: p = max(x, 0) into the state (MPS_CPU.cpp:715-743)
static int press_corr(mps_ctx* c)
{
    PpeBuffers* B = c->ppe;
    if (c->prec) launch_press_corr<double>(B->n, B->x, (double*)c->f[c->aux][6], c->st);
    else launch_press_corr<float>(B->n, B->x, (float*)c->f[c->aux][6], c->st);
    HIPOK(hipGetLastError());
    return MPS_OK;
}

extern "C" int mps_ppe_press_corr(mps_ctx* c)
{
    if (!c || !c->ppe || !c->ppe->assembled) return MPS_ERR_STATE;
    int rc = press_corr(c);
    if (rc) return rc;
    HIPOK(hipStreamSynchronize(c->st));
    return MPS_OK;
}

extern "C" int mps_stage_pnd(mps_ctx* c)
{
    if (!c) return MPS_ERR_ARG;
    int rc = mps_ctx_stage(c, 5);
    if (rc) return rc;
    HIPOK(hipStreamSynchronize(c->st));
    return MPS_OK;
}

int mps_mlp_guess(mps_ctx* c);      // mps_mlp.hip: features + inference into the guess buffer
int mps_mlp_prepare(mps_ctx* c);    // mps_mlp.hip: features of the assembled system
int mps_mlp_record(mps_ctx* c);     // mps_mlp.hip: append (features, solution) to the dataset
int mps_mlp_recording(mps_ctx* c);

// MPS_CPU_EXPL::step (MPS_CAL_v4.0/MPS_CPU_EXPL.cpp:79-135), two-pass operators:
// dt, [re-bin], viscosity + gravity, collision, n*, Poisson build + CG solve,
// p = max(x, 0), pressure-gradient correction (type-2 j skipped as in
// MPS_CPU::d3_GradPress).
extern "C" int mps_step_semi(mps_ctx* c, int nsteps, double alpha, int x0_kind, double tol,
                             int* iters_total, double* solve_ms)
{
    if (!c || nsteps < 0 || x0_kind < 0 || x0_kind > 3) return MPS_ERR_ARG;
    const bool every = c->P.rebin_every_move != 0;
    const int keep_skip = c->kc.gradp_skip_dummy;
    int total = 0;
    double ms = 0.0;
    int rc = MPS_OK;
    for (int s = 0; s < nsteps && rc == MPS_OK; s++) {
        launch_dt(c->kc, c->sc, c->st);
        if ((rc = mps_ctx_motion(c))) break;
        if (!every && (rc = mps_ctx_divide(c))) break;
        if ((rc = mps_ctx_stage(c, 0))) break;
        if (every && (rc = mps_ctx_divide(c))) break;
        if ((rc = mps_ctx_stage(c, 1))) break;
        if (every && (rc = mps_ctx_divide(c))) break;
        if ((rc = mps_ctx_stage(c, 5))) break;
        long nnz = 0;
        if ((rc = mps_ppe_assemble(c, alpha, &nnz))) break;
        const bool rec = mps_mlp_recording(c) != 0;
        if (x0_kind == 3) {
            if ((rc = mps_mlp_guess(c))) break;
        } else if (rec && (rc = mps_mlp_prepare(c))) {
            break;
        }
        PpeBuffers* B = c->ppe;
        HIPOK(hipEventRecord(B->t0, c->st));
        int it = 0;
        double rel = 0.0;
        rc = mps_ppe_solve(c, 1, x0_kind == 2 ? 1 : x0_kind, nullptr, tol, 100000, &it, &rel);
        HIPOK(hipEventRecord(B->t1, c->st));
        HIPOK(hipEventSynchronize(B->t1));
        float t = 0.0f;
        HIPOK(hipEventElapsedTime(&t, B->t0, B->t1));
        ms += t;
        total += it;
        if (rc) break;
        if (rec && (rc = mps_mlp_record(c))) break;
        if ((rc = press_corr(c))) break;
        c->kc.gradp_skip_dummy = 1;
        rc = mps_ctx_stage(c, 4);
        c->kc.gradp_skip_dummy = keep_skip;
        if (rc) break;
        if (every && (rc = mps_ctx_divide(c))) break;
        c->steps++;
    }
    if (iters_total) *iters_total = total;
    if (solve_ms) *solve_ms = ms;
    if (rc) return rc;
    return mps_sync(c);
}
