/*
 * mps_amd.h -- C ABI of the MI355X (gfx950) MPS particle time step.
 *
 * Plain C: opaque context, plain pointers and sizes, int status codes, no
 * exit().  Host buffers are caller-owned; device buffers are owned by the
 * context.  A context is not thread-safe.  Each entry point cites the
 * reference interface it replaces (paths relative to the dlut414/MPS_SOLVERS
 * root).
 */
#ifndef MPS_AMD_H
#define MPS_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mps_ctx mps_ctx;

/* Status codes.  The reference exits the process instead
 * (exit(3) missing GEO.dat, exit(5) thin grid, exit(1) dt too small). */
enum {
    MPS_OK = 0,
    MPS_ERR_ARG = 1,     /* bad argument */
    MPS_ERR_IO = 3,      /* file missing or malformed (MPS_CAL_expl_v1.0/MPS.cpp:229-234) */
    MPS_ERR_GRID = 5,    /* cells too thin (MPS_CAL_expl_v1.0/MPS.cpp:506-515) */
    MPS_ERR_DT = 6,      /* dt fell below dt_min (MPS_CAL_expl_v1.0/MPS_GPU.cpp:770-776) */
    MPS_ERR_HIP = 7,     /* HIP runtime error */
    MPS_ERR_NODEV = 8,   /* no gfx950 device */
    MPS_ERR_CAP = 9,     /* caller buffer too small */
    MPS_ERR_STATE = 10   /* call not valid in the current state */
};

/* Run parameters: the 16 entries of input.dat (MPS_CAL_expl_v1.0/MPS.cpp:62-223)
 * plus the build's precision and step-order choices.  R0, RLAP and the
 * collision distance are multiples of dp, dt_min_mult a multiple of dt_max,
 * as in input.dat; they are multiplied in the state precision. */
typedef struct mps_params {
    int dim;               /* 2 or 3 */
    int precision;         /* 0: FP32 state, FP64 n and (n-n0)/n0; 1: FP64 state */
    int gradp_skip_dummy;  /* 1: gradient sum skips type-2 j (MPS_CAL_v4.0/MPS_CPU.cpp:430) */
    int rebin_every_move;  /* 0: one re-binning per step (MPS_CAL_/MPS_GPU.cpp:611-714);
                              1: after every position update (MPS_CAL_expl_v1.0/MPS_GPU_EXPL.cpp:85-147) */
    double rho, nu, dp;
    double re_n_mult, re_lap_mult;
    double beta, cs, cfl;
    double col_mult, col_rate;
    double dt_max, dt_min_mult;
    double tt, tout, max_dis;
    double g[3];
} mps_params;

/* ---------------------------------------------------------------- scene I/O */

/* Defaults of MPS_CAL_expl_v1.0/input.dat with g = (0, 0, -9.81). */
int mps_params_default(mps_params* p);
/* MPS::loadCase (MPS_CAL_expl_v1.0/MPS.cpp:62-223): label line / value line pairs. */
int mps_load_case(const char* path, mps_params* p);
int mps_write_case(const char* path, const mps_params* p);
/* MPS::createGeo (MPS_CAL_expl_v1.0/MPS.cpp:225-300): "np nb2 nb1", then all
 * positions, all velocities, all pressures.  Order: type 2, type 1, fluid.
 * Call with pos == NULL to query the counts.  as_float != 0 parses each
 * number to the nearest float, as the float build's %f does. */
int mps_geo_read(const char* path, int as_float, int* np, int* nb2, int* nb1, double* pos,
                 double* vel, double* press);
int mps_geo_write(const char* path, int np, int nb2, int nb1, const double* pos, const double* vel,
                  const double* press);
/* Seeded cubic-lattice dam break: fluid block fx*fy*fz lattice sites in the
 * (0,0,0) corner of an open-top tank of nx*ny*nz inner sites, lined with
 * layers_wall type-1 and layers_dummy type-2 layers; uniform jitter of
 * +-jitter*dp per coordinate from a counter-based hash of (seed, index).
 * 2-D when dim == 2 (y = 0).  Order as in GEO.dat (MPS_INI/main.cpp:64-94).
 * Call with pos == NULL to query the counts. */
int mps_scene_dambreak(int dim, int fx, int fy, int fz, int nx, int ny, int nz, int layers_wall,
                       int layers_dummy, double dp, double jitter, unsigned long long seed,
                       int* np, int* nb2, int* nb1, double* pos);
/* Seeded 3-D wave tank (no reference generator; the reference's MOTION
 * class, MPS_CAL_/MOTION.cpp:11-16, is an empty stub): water nx*ny*depth in an
 * open-top tank of nx*ny*nz inner sites whose x = 0 wall is a piston that may
 * travel +-travel sites; floor and side walls reach travel sites behind it.
 * piston (np entries, may be NULL) flags the piston's wall and dummy layers.
 * Same layering, jitter and block order as mps_scene_dambreak. */
int mps_scene_wavetank(int nx, int ny, int nz, int depth, int travel, int layers_wall, int layers_dummy,
                       double dp, double jitter, unsigned long long seed, int* np, int* nb2, int* nb1,
                       double* pos, unsigned char* piston);

/* ------------------------------------------------------------------ context */

/* MPS_GPU::Initial (MPS_CAL_expl_v1.0/MPS_GPU.cpp:63-115): cell grid from the
 * initial bounding box, first re-binning, then calOnce (n0, lambda, normals).
 * pos/vel are interleaved xyz (3*np), press has np entries; types follow the
 * index ranges of GEO.dat. */
int mps_create(const mps_params* p, int np, int nb2, int nb1, const double* pos, const double* vel,
               const double* press, mps_ctx** out);
int mps_destroy(mps_ctx* c);

/* MPS_GPU_EXPL::step (MPS_CAL_expl_v1.0/MPS_GPU_EXPL.cpp:85-147 /
 * MPS_CAL_/MPS_GPU.cpp:611-714), nsteps times, enqueued on the context's
 * stream; returns after the last step completes. */
int mps_step(mps_ctx* c, int nsteps);
/* Same, but returns without waiting (errors surface at the next mps_sync). */
int mps_step_async(mps_ctx* c, int nsteps);
int mps_sync(mps_ctx* c);

/* Getters of MPS_CAL_/MPS_GPU.h:158-166 (getNp, getPos, getVel, getPres,
 * getId, getType).  order 0: current cell-sorted order; 1: by particle id.
 * Any output pointer may be NULL.  pos/vel interleaved xyz. */
int mps_get_np(mps_ctx* c, int* np_total, int* np_active);
int mps_get_state(mps_ctx* c, int order, double* pos, double* vel, double* press, double* n,
                  int* id, int* type);
/* Overwrite position / velocity / pressure in the current (sorted) order.
 * New positions may be anywhere, so the next re-binning sorts from scratch. */
int mps_set_state(mps_ctx* c, const double* pos, const double* vel, const double* press);
/* out[0..9] = n0, lambda, 2d/(n0 lambda), dt, time, ncells, collision hits of
 * the last collision sweep, excluded particles, cell size, step count. */
int mps_get_scalars(mps_ctx* c, double* out10);
/* geom[0..3] = left, back, bottom, h; dims[0..3] = dx, dy, dz, ncells */
int mps_get_grid(mps_ctx* c, double* geom4, int* dims4);
/* MPS_GPU::writeCase (MPS_CAL_expl_v1.0/MPS_GPU.cpp:138-152). */
int mps_write_out(mps_ctx* c, const char* path);
/* The legacy-VTK particle file of MPS_POST (MPS_POST/MPS_POST.cpp:93-172:
 * WriteHead, WritePoint, WriteCell1, WriteAttribute) for the current state,
 * particles in id order. */
int mps_write_vtk(mps_ctx* c, const char* path);

/* Prescribed motion (the role of MOTION::doMotion, an empty stub in
 * MPS_CAL_/MOTION.cpp:11-16): the particles flagged in moving (np entries,
 * input order = particle id) follow x_axis(t) = x_axis(now) + A (sin(w t) -
 * sin(w now)), w = 2 pi freq, with velocity A w cos(w t) along axis and zero
 * across it.  Applied at the start of every step, right after dt, for t =
 * the step's end time, before any sweep.  amplitude 0 or moving == NULL
 * switches it off. */
int mps_set_motion(mps_ctx* c, const unsigned char* moving, int axis, double amplitude, double freq);

/* ---------------------------------------------- per-stage entry points (parity) */

/* MPS_GPU::divideCell + sort_* + sort_normal (MPS_CAL_expl_v1.0/MPS_GPU.cpp:248-373).
 * perm (np_total, new -> old) and cell_start (ncells + 1) may be NULL.  The
 * permutation is the stable sort of the current order by cell, as the
 * reference's, whichever way it is formed (a radix sort of a host-supplied
 * state, or a gather from the previous cells of a state this library sorted). */
int mps_stage_sort(mps_ctx* c, int* perm, int* cell_start);
/* MPS_GPU_EXPL::calVisc_expl (MPS_CAL_expl_v1.0/MPS_GPU_EXPL.cpp:152-194) */
int mps_stage_visc(mps_ctx* c);
/* MPS_GPU::collision (MPS_CAL_expl_v1.0/MPS_GPU.cpp:892-986) */
int mps_stage_collision(mps_ctx* c);
/* MPS_GPU::cal_n + MPS_GPU_EXPL::calPres_expl (MPS_GPU.cpp:830-887, MPS_GPU_EXPL.cpp:199-238) */
int mps_stage_pnd_pres(mps_ctx* c);
/* MPS_GPU::calDash (MPS_CAL_expl_v1.0/MPS_GPU.cpp:782-825) */
int mps_stage_gradp(mps_ctx* c);
/* MPS_GPU::update_dt (MPS_CAL_expl_v1.0/MPS_GPU.cpp:761-777) */
int mps_stage_dt(mps_ctx* c);
/* Fix dt for the following stage calls (no reference counterpart: lets a
 * parity test pin dt). */
int mps_set_dt(mps_ctx* c, double dt);
/* Within-cutoff neighbour lists (which 0: R0, 1: RLAP) of the current sorted
 * state, columns in the reference visiting order.  row_ptr has np_active + 1
 * entries; *nnz receives the total; MPS_ERR_CAP when cap is too small. */
int mps_neighbor_list(mps_ctx* c, int which, int* row_ptr, int* col, long cap, long* nnz);
/* Accumulated per-stage device milliseconds since the last reset:
 * ms[0..7] = dt, sort, visc, collision, pnd+pres, dummy p, gradp, total step. */
int mps_stage_times(mps_ctx* c, double* ms8, int reset);
int mps_enable_timing(mps_ctx* c, int on);

/* ------------------------------------------------ pressure Poisson (semi-implicit) */

/* MPS_GPU::cal_n alone (MPS_CAL_v4.0/MPS_CPU.cpp:573-614): n* of the current state. */
int mps_stage_pnd(mps_ctx* c);
/* MPS_CPU::buildPoisson (MPS_CAL_v4.0/MPS_CPU.cpp:619-710) as CSR on the device,
 * from the current n, positions and dt.  alpha is the compressibility of the
 * diagonal term (1/(rho c^2) in the reference; 0 gives the textbook form). */
int mps_ppe_assemble(mps_ctx* c, double alpha, long* nnz);
int mps_ppe_info(mps_ctx* c, int* nrows, long* nnz);
/* Copy the assembled system: row_ptr (nrows + 1), col / val (nnz, FP64 values),
 * b, current iterate x, Dirichlet flags bc (nrows).  Any pointer may be NULL. */
int mps_ppe_get(mps_ctx* c, int* row_ptr, int* col, double* val, double* b, double* x,
                unsigned char* bc);
int mps_ppe_get_x(mps_ctx* c, double* x);
/* MPS_CPU::DoCG (MPS_CAL_v4.0/MPS_CPU.cpp:829-888) / cudaker::CG
 * (MPS_CAL_v4.0/cuda_ker.cu:171-297): solve A x = b to |b - A x| <= tol |b|
 * (FP64 residual).  mixed 1: FP32-SpMV CG inside FP64 iterative refinement;
 * 0: FP64 CG.  x0_kind 0: zero, 1: current pressure (previous step),
 * 2: host array x0, 3: the learned guess left by mps_mlp_infer.
 * Returns MPS_ERR_STATE when maxit inner iterations did not reach tol. */
int mps_ppe_solve(mps_ctx* c, int mixed, int x0_kind, const double* x0, double tol, int maxit,
                  int* iters, double* relres);
/* MPS_CPU::pressCorr (MPS_CAL_v4.0/MPS_CPU.cpp:734-743): p = max(x, 0) into the state. */
int mps_ppe_press_corr(mps_ctx* c);
/* MPS_CPU_EXPL::step (MPS_CAL_v4.0/MPS_CPU_EXPL.cpp:79-135): dt, viscosity,
 * collision, n*, Poisson build + mixed-precision CG to tol, pressCorr,
 * pressure-gradient correction.  Reports total CG iterations and device
 * milliseconds spent in the solves. */
int mps_step_semi(mps_ctx* c, int nsteps, double alpha, int x0_kind, double tol, int* iters_total,
                  double* solve_ms);

/* --------------------------------------------- learned CG warm start (no reference counterpart) */

/* Per-particle MLP 16 -> 64 -> 64 -> 64 -> 1 (SiLU), BF16 MFMA with FP32
 * accumulation, FP32 master parameters (9,473 values, flat order W1[16x64],
 * b1, W2[64x64], b2, W3[64x64], b3, W4[64], b4).  The guess is
 * x0 = p_prev + P_s * y with P_s = rho |g| 64 dp.  These entry points have no
 * counterpart in the reference; they feed mps_ppe_solve(x0_kind = 3). */
int mps_mlp_nparams(void);
int mps_mlp_init(mps_ctx* c, unsigned long long seed);
int mps_mlp_set_params(mps_ctx* c, const float* theta);
int mps_mlp_get_params(mps_ctx* c, float* theta);
/* y = MLP(X) for n caller feature rows (n x 16). */
int mps_mlp_forward(mps_ctx* c, const float* X, int n, float* y);
/* Gradient (9,473) and value of mean((MLP(X) - t)^2) over n caller samples. */
int mps_mlp_grad(mps_ctx* c, const float* X, const float* t, int n, float* grad, double* loss);
/* 16 features per row of the assembled Poisson system (F may be NULL). */
int mps_mlp_features(mps_ctx* c, float* F, int* nrows);
/* Features + inference into the solver's guess buffer. */
int mps_mlp_infer(mps_ctx* c);
/* When on, mps_step_semi appends (features, converged correction) of every
 * solve to a device-resident training set. */
int mps_mlp_set_record(mps_ctx* c, int on);
int mps_mlp_dataset(mps_ctx* c, long* count, float* X, float* T);
int mps_mlp_dataset_append(mps_ctx* c, const float* X, const float* T, long n);
int mps_mlp_dataset_clear(mps_ctx* c);
/* iters Adam steps (lr) on minibatches of the training set. */
int mps_mlp_train(mps_ctx* c, int iters, int batch, double lr, unsigned long long seed,
                  double* loss_first, double* loss_last);

#ifdef __cplusplus
}
#endif
#endif
