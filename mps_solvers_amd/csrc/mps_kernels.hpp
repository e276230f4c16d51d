// mps_kernels.hpp -- host-visible launchers for the gfx950 MPS kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mps_device.hpp"

namespace mps {

// Device pointers of one copy of the particle state (structure of arrays).
template <typename Real>
struct StateView {
    Real *x, *y, *z, *u, *v, *w, *p;
    double* n;
    int *type, *id, *normal;
};

// Scalars that stay on the device between kernels (no host round trip per step).
struct DevScalars {
    double dt;          // time step of the current step
    double time;        // simulated time after the current step
    double n0, lambda, coef;
    unsigned int maxu2_bits;   // max |u|^2 of fluid particles (float bits), accumulated by the gradient sweep
    int err;            // 1: dt fell below dt_min
    int nexcl;          // particles that left the grid
    int argmax;
    unsigned long long nmax_bits;
    int ncol;           // collision hits in the last collision sweep
    int pad_;
};

struct StepConst {
    double rho, nu, cs, cfl, dp, dt_max, dt_min, col_rate;
    double g[3];
    double umax2_static;   // max |u|^2 over particles that never move
    double mov_amp, mov_omega;   // prescribed motion (amplitude 0: none); its speed enters dt
    int dim;
    int gradp_skip_dummy;
    int pad_;
};

namespace wv {
template <typename Real>
void launch_pnd(const Grid& g, const int* cs, StateView<Real> s, const Thresh<Real>& th, const StepConst& k,
                DevScalars* sc, int all_types, int with_pressure, hipStream_t st);
template <typename Real>
void launch_visc(const Grid& g, const int* cs, StateView<Real> a, StateView<Real> b, const Thresh<Real>& th,
                 const StepConst& k, DevScalars* sc, hipStream_t st);
template <typename Real>
void launch_collision(const Grid& g, const int* cs, StateView<Real> a, StateView<Real> b, const Thresh<Real>& th,
                      const StepConst& k, DevScalars* sc, hipStream_t st);
template <typename Real>
void launch_gradp(const Grid& g, const int* cs, StateView<Real> a, StateView<Real> b, const Thresh<Real>& th,
                  const StepConst& k, DevScalars* sc, hipStream_t st);
}  // namespace wv

// sort
int sort_scratch_words(int n);
// hash (and, with cnt, per-cell key counts: cnt[0..ncells] must hold zeros)
template <typename Real>
void launch_hash(const Grid& g, int n, StateView<Real> s, uint32_t* key, uint32_t* val, int* nexcl,
                 uint32_t* cnt, hipStream_t st);
void launch_radix_sort(int n, int key_bits, uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1,
                       uint32_t* hist, uint32_t** kres, uint32_t** vres, hipStream_t st);
void launch_cell_bounds(const uint32_t* keys, int n, int ncells, int* cs, hipStream_t st);
// stable re-binning of a previously sorted order: cs_new holds the hash's
// counts on entry and the new cell starts on return; perm receives the
// permutation; tiles needs scan_tile_words(ncells + 1) words
int scan_tile_words(int m);
void launch_gather_sort(const Grid& g, int n, const uint32_t* key, const int* cs_old, uint32_t* cs_new,
                        uint32_t* tiles, uint32_t* perm, hipStream_t st);
// gather the state into sorted order; also clears zero[0..nzero)
template <typename Real>
void launch_reorder(int n, const uint32_t* perm, StateView<Real> src, StateView<Real> dst, int* inv,
                    int* normal_tmp, int* zero, int nzero, hipStream_t st);
// copy positions and velocities of the excluded slots [cs[ncells], n) from one
// state buffer into the other, so that both hold them until the next sort
template <typename Real>
void launch_tail_copy(int n, const int* np_active, StateView<Real> from, StateView<Real> to, hipStream_t st);

// sweeps
template <typename Real>
void launch_pnd(const Grid& g, const int* cs, StateView<Real> s, const Thresh<Real>& th,
                const StepConst& k, DevScalars* sc, int all_types, int with_pressure, hipStream_t st);
template <typename Real>
void launch_dummy_p(int n, StateView<Real> s, hipStream_t st);
template <typename Real>
void launch_visc(const Grid& g, const int* cs, StateView<Real> a, StateView<Real> b,
                 const Thresh<Real>& th, const StepConst& k, DevScalars* sc, hipStream_t st);
template <typename Real>
void launch_collision(const Grid& g, const int* cs, StateView<Real> a, StateView<Real> b,
                      const Thresh<Real>& th, const StepConst& k, DevScalars* sc, hipStream_t st);
template <typename Real>
void launch_gradp(const Grid& g, const int* cs, StateView<Real> a, StateView<Real> b,
                  const Thresh<Real>& th, const StepConst& k, DevScalars* sc, hipStream_t st);
void launch_dt(const StepConst& k, DevScalars* sc, hipStream_t st);
template <typename Real>
void launch_calonce(const Grid& g, const int* cs, int np, StateView<Real> s, const Thresh<Real>& th,
                    const StepConst& k, DevScalars* sc, hipStream_t st);
template <typename Real>
void launch_nlist(const Grid& g, const int* cs, int np, StateView<Real> s, const Thresh<Real>& th,
                  int which, int* row_ptr, int* col, int fill, hipStream_t st);

// prescribed wall motion (mps_motion.hip)
template <typename Real>
void launch_motion_ref(int n, StateView<Real> s, const unsigned char* mov, double* xref, int axis, double amp,
                       double omega, const DevScalars* sc, hipStream_t st);
template <typename Real>
void launch_motion(int n, StateView<Real> s, const unsigned char* mov, const double* xref, int axis, double amp,
                   double omega, const DevScalars* sc, hipStream_t st);

}  // namespace mps
