/*
 * mps_oracle.h -- CPU oracle for the MPS particle time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * solver's operators (dlut414/MPS_SOLVERS, trees MPS_CAL_expl_v1.0 and
 * MPS_CAL_v4.0), used by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py as the checker.  The product library
 * (mps_solvers_amd) never links, loads or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Sums are accumulated in double precision (the "FP64 oracle").  The
 * integer/neighbour decisions (cell hash, cutoff tests) can be evaluated
 * either in float arithmetic (f32 = 1, MPS_CAL_expl_v1.0 where real = float,
 * typedef.h:20) or in double arithmetic (f32 = 0, MPS_CAL_v4.0 where
 * real = double, typedef.h:20), so that neighbour lists and sort
 * permutations can be compared bit-exactly against either precision.
 */
#ifndef MPS_ORACLE_H
#define MPS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    int dim;              /* 2 or 3 */
    int f32;              /* 1: hash + cutoff tests in float arithmetic */
    int gradp_skip_dummy; /* 1: pressure-gradient sum skips type-2 j (MPS_CAL_v4.0/MPS_CPU.cpp:430) */
    int pad_;
    double rho, nu, dp;
    double re_n_mult, re_lap_mult;   /* R0, RLAP in units of dp (MPS.cpp:113,123) */
    double beta, cs, cfl;
    double col_mult, col_rate;       /* collision distance in units of dp (MPS.cpp:160) */
    double dt_max, dt_min_mult;      /* MIN TIME STEP is a multiple of MAX (MPS.cpp:188) */
    double max_dis;                  /* bbox margin (MPS.cpp:479-484) */
    double g[3];
} orc_params;

typedef struct orc_grid {
    double left, back, bottom, h;    /* float-representable in f32 mode */
    int dx, dy, dz, sheet, ncells, pad_;
} orc_grid;

/* Effective thresholds: out[0]=re_n, [1]=re_lap, [2]=col_dis, [3]=re_n^2 (p-hat test),
 * [4]=dt_min, [5]=cell size.  Products are formed in the mode's precision. */
void orc_thresholds(const orc_params* P, double out[6]);

int  orc_grid_init(const orc_params* P, int np, const double* x, const double* y,
                   const double* z, orc_grid* G);
int  orc_hash(const orc_params* P, const orc_grid* G, int np, const double* x,
              const double* y, const double* z, const int* type, int* key);
void orc_sort(int np, int ncells, const int* key, int* perm, int* index, int* cell_start);

int  orc_neighbors(const orc_params* P, const orc_grid* G, int np, const double* x,
                   const double* y, const double* z, const int* cell_start, int which,
                   int* row_ptr, int* col, long cap);

void orc_calonce(const orc_params* P, const orc_grid* G, int np, const double* x,
                 const double* y, const double* z, const int* type, const int* cell_start,
                 double* n, double* n0_out, double* lambda_out, int* argmax_out, int* normal);

void orc_pnd(const orc_params* P, const orc_grid* G, int np, const double* x, const double* y,
             const double* z, const int* type, const int* cell_start, double* n);
void orc_pres_expl(const orc_params* P, int np, const double* n, double n0, const int* type,
                   const int* normal, double* p);
void orc_visc(const orc_params* P, const orc_grid* G, int np, const double* x, const double* y,
              const double* z, const double* u, const double* v, const double* w,
              const int* type, const int* cell_start, double coef, double dt,
              double* xo, double* yo, double* zo, double* uo, double* vo, double* wo);
int  orc_collision(const orc_params* P, const orc_grid* G, int np, const double* x,
                   const double* y, const double* z, const double* u, const double* v,
                   const double* w, const int* type, const int* cell_start, double dt,
                   double* xo, double* yo, double* zo, double* uo, double* vo, double* wo);
void orc_lapvel(const orc_params* P, const orc_grid* G, int np, const double* x, const double* y,
                const double* z, const double* u, const double* v, const double* w,
                const int* cell_start, double coef, double* lx, double* ly, double* lz);
void orc_gradp_vec(const orc_params* P, const orc_grid* G, int np, const double* x,
                   const double* y, const double* z, const double* p, const int* type,
                   const int* cell_start, double n0, double* gx, double* gy, double* gz);
void orc_gradp(const orc_params* P, const orc_grid* G, int np, const double* x, const double* y,
               const double* z, const double* u, const double* v, const double* w,
               const double* p, const int* type, const int* cell_start, double n0, double dt,
               double* xo, double* yo, double* zo, double* uo, double* vo, double* wo);
double orc_update_dt(const orc_params* P, int np, const double* u, const double* v,
                     const double* w, int* too_small);

long orc_build_poisson(const orc_params* P, const orc_grid* G, int np, const double* x,
                       const double* y, const double* z, const int* type, const int* cell_start,
                       const double* n, const double* p_prev, double n0, double lambda,
                       double dt, double alpha, int* row_ptr, int* col, double* val,
                       double* b, double* x0, unsigned char* bc, long cap);
int  orc_cg(int np, const int* row_ptr, const int* col, const double* val, const double* b,
            double* x, double tol, int relative, int maxit, double* rr_out);
void orc_spmv(int np, const int* row_ptr, const int* col, const double* val, const double* x,
              double* y);

/* Whole-step simulator (used as the CPU baseline and for trajectory parity). */
typedef struct orc_sim orc_sim;
orc_sim* orc_sim_create(const orc_params* P, int np, int nb2, int nb1, const double* x,
                        const double* y, const double* z, const double* u, const double* v,
                        const double* w, const double* p);
void orc_sim_destroy(orc_sim* S);
int  orc_sim_step(orc_sim* S, int nsteps, int rebin_every_move);
/* Semi-implicit steps (MPS_CPU_EXPL::step order, two-pass, FP64 CG to a relative tol). */
int  orc_sim_step_semi(orc_sim* S, int nsteps, int rebin_every_move, double alpha, double tol, int* iters);
int  orc_sim_np(const orc_sim* S);
void orc_sim_get(const orc_sim* S, double* x, double* y, double* z, double* u, double* v,
                 double* w, double* p, double* n, int* type, int* id);
void orc_sim_consts(const orc_sim* S, double* out /* n0, lambda, dt, time, coef, ncells */);
void orc_sim_grid(const orc_sim* S, orc_grid* G);
/* Prescribed sinusoidal motion of the particles flagged by id (NULL: off). */
void orc_sim_set_motion(orc_sim* S, const unsigned char* moving_by_id, int axis, double amp, double freq);

#ifdef __cplusplus
}
#endif
#endif
