// mps_ctx.hpp -- the device context behind the C ABI (private to the library).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/mps_amd.h"
#include "mps_kernels.hpp"

enum { T_DT, T_SORT, T_VISC, T_COL, T_PND, T_DUMMY, T_GRAD, T_STEP, T_N };

struct Thr {
    mps::Thresh<float> f;
    mps::Thresh<double> d;
};

struct PpeBuffers;   // mps_ppe_capi.hip
struct MlpBuffers;   // mps_mlp_capi

struct mps_ctx {
    mps_params P;
    int prec = 0;
    int nt = 0, nb2 = 0, nb1 = 0;
    mps::Grid g{};
    Thr th{};
    mps::StepConst kc{};
    hipStream_t st = nullptr;
    int key_bits = 0;
    // two copies of the state; kin selects x..w, aux selects p, n, type, id, normal
    void* f[2][7] = {};
    double* n[2] = {};
    int *type[2] = {}, *id[2] = {}, *normal[2] = {};
    int kin = 0, aux = 0;
    uint32_t *k0 = nullptr, *v0 = nullptr, *k1 = nullptr, *v1 = nullptr, *hist = nullptr;
    uint32_t* perm = nullptr;   // last sort's permutation (one of v0/v1)
    int *inv = nullptr, *ntmp = nullptr, *cs = nullptr;
    // the other cell-start array (all zeros between sorts: the next re-binning
    // counts keys into it), the tile totals of its scan, and whether the
    // current order is a previous sort's (re-bin by gathering) or not (radix sort)
    int* cs_next = nullptr;
    uint32_t* cs_tiles = nullptr;
    bool binned = false;
    mps::DevScalars* sc = nullptr;
    long steps = 0;
    bool timing = false;
    hipEvent_t ev[T_N][2] = {};
    double ms[T_N] = {};
    std::vector<void*> allocs;
    // prescribed motion: flags and reference coordinate by particle id
    unsigned char* mov = nullptr;
    double* xref = nullptr;
    int mov_axis = 0;   // amplitude and angular frequency live in kc
    PpeBuffers* ppe = nullptr;
    MlpBuffers* mlp = nullptr;

    template <typename Real>
    mps::StateView<Real> view(int k, int a)
    {
        mps::StateView<Real> s;
        s.x = (Real*)f[k][0]; s.y = (Real*)f[k][1]; s.z = (Real*)f[k][2];
        s.u = (Real*)f[k][3]; s.v = (Real*)f[k][4]; s.w = (Real*)f[k][5];
        s.p = (Real*)f[a][6]; s.n = n[a]; s.type = type[a]; s.id = id[a]; s.normal = normal[a];
        return s;
    }
    size_t rsize() const { return prec ? sizeof(double) : sizeof(float); }
};


int mps_ctx_divide(mps_ctx* c);
int mps_ctx_motion(mps_ctx* c);   // prescribed motion for the step about to run (no-op when off)
int mps_ctx_stage(mps_ctx* c, int stage);   // 0 visc, 1 collision, 2 pnd+p, 3 dummy p, 4 grad p, 5 pnd only
void mps_ppe_free(mps_ctx* c);
void mps_mlp_free(mps_ctx* c);
