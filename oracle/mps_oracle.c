/*
 * mps_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY) for the MPS particle time step.
 *
 * Clean-room restatement of the reference operators.  Only tests/, the smoke
 * check and the CPU-baseline leg of bench.py use it; the product library
 * never does.  Reference citations are relative to the reference repository
 * root (dlut414/MPS_SOLVERS).
 *
 * Conventions shared with the reference:
 *   - particles are kept in cell-sorted order; cell c owns [cell_start[c], cell_start[c+1]);
 *   - the 27-cell stencil is visited in (dz, dy, dx) = -1..1 order, i.e. in
 *     ascending cell index (MPS_CAL_expl_v1.0/MPS_GPU.cpp:193-240), so
 *     candidates j are visited in ascending sorted index;
 *   - w(r; re) = re / r - 1 for r < re, else 0 (MPS_CAL_expl_v1.0/MPS.cpp:301-306).
 * Unlike the reference's link table, the stencil never wraps across rows:
 * wrapped cells lie more than one cell size away, so the within-cutoff
 * neighbour sets are the same.
 */
#include "mps_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* thresholds and geometry                                              */
/* ------------------------------------------------------------------ */

/* Thresholds as MPS::loadCase forms them: R0, RLAP and COLLISION DISTANCE are
 * read as multiples of DP and multiplied in `real` precision
 * (MPS_CAL_expl_v1.0/MPS.cpp:113,123,160,188). */
void orc_thresholds(const orc_params* P, double out[6])
{
    if (P->f32) {
        float dp = (float)P->dp;
        float rn = (float)P->re_n_mult * dp;
        float rl = (float)P->re_lap_mult * dp;
        float cd = (float)P->col_mult * dp;
        float rn2 = rn * rn;
        float dtm = (float)P->dt_min_mult * (float)P->dt_max;
        out[0] = rn; out[1] = rl; out[2] = cd; out[3] = rn2; out[4] = dtm;
        out[5] = rn > rl ? rn : rl;
    } else {
        double dp = P->dp;
        double rn = P->re_n_mult * dp, rl = P->re_lap_mult * dp;
        out[0] = rn; out[1] = rl; out[2] = P->col_mult * dp; out[3] = rn * rn;
        out[4] = P->dt_min_mult * P->dt_max; out[5] = rn > rl ? rn : rl;
    }
}

/* MPS::allocateCellList (MPS_CAL_expl_v1.0/MPS.cpp:456-531): bounding box of
 * the initial positions widened by max_dis, cell size = max(R0, RLAP),
 * cells per axis = floor(L / h) + 1; 2-D collapses y. Returns 5 when the
 * grid is too thin (the reference exit(5)). */
int orc_grid_init(const orc_params* P, int np, const double* x, const double* y,
                  const double* z, orc_grid* G)
{
    double th[6];
    orc_thresholds(P, th);
    memset(G, 0, sizeof(*G));
    if (np <= 0) return 3;
    if (P->f32) {
        float l = (float)x[0], r = l, bk = (float)y[0], fr = bk, bo = (float)z[0], tp = bo;
        for (int i = 0; i < np; i++) {
            float a = (float)x[i], b = (float)y[i], c = (float)z[i];
            if (a < l) l = a; if (a > r) r = a;
            if (b < bk) bk = b; if (b > fr) fr = b;
            if (c < bo) bo = c; if (c > tp) tp = c;
        }
        float md = (float)P->max_dis;
        l -= md; r += md; bk -= md; fr += md; bo -= md; tp += md;
        if (P->dim == 2) { bk = 0.0f; fr = 0.0f; }
        float h = (float)th[5];
        float Lx = r - l, Ly = fr - bk, Lz = tp - bo;
        G->dx = (int)(unsigned)(Lx / h) + 1;
        G->dy = (int)(unsigned)(Ly / h) + 1;
        G->dz = (int)(unsigned)(Lz / h) + 1;
        G->left = l; G->back = bk; G->bottom = bo; G->h = h;
    } else {
        double l = x[0], r = l, bk = y[0], fr = bk, bo = z[0], tp = bo;
        for (int i = 0; i < np; i++) {
            if (x[i] < l) l = x[i]; if (x[i] > r) r = x[i];
            if (y[i] < bk) bk = y[i]; if (y[i] > fr) fr = y[i];
            if (z[i] < bo) bo = z[i]; if (z[i] > tp) tp = z[i];
        }
        double md = P->max_dis;
        l -= md; r += md; bk -= md; fr += md; bo -= md; tp += md;
        if (P->dim == 2) { bk = 0.0; fr = 0.0; }
        double h = th[5];
        G->dx = (int)(unsigned)((r - l) / h) + 1;
        G->dy = (int)(unsigned)((fr - bk) / h) + 1;
        G->dz = (int)(unsigned)((tp - bo) / h) + 1;
        G->left = l; G->back = bk; G->bottom = bo; G->h = h;
    }
    G->sheet = G->dx * G->dy;
    G->ncells = G->sheet * G->dz;
    if (G->dx <= 2 || G->dz <= 2) return 5;
    if (G->dy <= 2 && P->dim == 3) return 5;
    return 0;
}

/* Cell hash of MPS_GPU::divideCell (MPS_CAL_expl_v1.0/MPS_GPU.cpp:259-288):
 * c_a = integer((pos_a - origin_a) / h), truncating toward zero; a particle
 * whose cell falls outside the grid is excluded (key = ncells).  Type-3
 * particles stay excluded.  Returns the number of newly excluded particles. */
int orc_hash(const orc_params* P, const orc_grid* G, int np, const double* x,
             const double* y, const double* z, const int* type, int* key)
{
    int excl = 0;
    for (int i = 0; i < np; i++) {
        if (type && type[i] == 3) { key[i] = G->ncells; continue; }
        double vx, vy, vz;
        if (P->f32) {
            float h = (float)G->h;
            vx = ((float)x[i] - (float)G->left) / h;
            vy = ((float)y[i] - (float)G->back) / h;
            vz = ((float)z[i] - (float)G->bottom) / h;
        } else {
            vx = (x[i] - G->left) / G->h;
            vy = (y[i] - G->back) / G->h;
            vz = (z[i] - G->bottom) / G->h;
        }
        double tx = trunc(vx), ty = trunc(vy), tz = trunc(vz);
        if (!(tx >= 0 && ty >= 0 && tz >= 0 && tx < G->dx && ty < G->dy && tz < G->dz)) {
            key[i] = G->ncells;
            excl++;
            continue;
        }
        key[i] = (int)tz * G->sheet + (int)ty * G->dx + (int)tx;
    }
    return excl;
}

/* Stable counting sort of MPS_GPU::divideCell (MPS_CAL_expl_v1.0/MPS_GPU.cpp:290-315):
 * exclusive scan of per-cell counts, then slots handed out in current order.
 * perm[new] = old, index[old] = new (the reference's i_index).  Excluded
 * particles (key == ncells) go after every cell in stable order; the
 * reference gives all of them slot np-1 (MPS_GPU.cpp:311-314), which loses
 * data when two leave in one step. */
void orc_sort(int np, int ncells, const int* key, int* perm, int* index, int* cell_start)
{
    int* cnt = (int*)calloc((size_t)ncells + 2, sizeof(int));
    for (int i = 0; i < np; i++) cnt[key[i]]++;
    int acc = 0;
    for (int c = 0; c <= ncells; c++) { int t = cnt[c]; cell_start[c] = acc; cnt[c] = acc; acc += t; }
    for (int i = 0; i < np; i++) {
        int s = cnt[key[i]]++;
        index[i] = s;
        perm[s] = i;
    }
    free(cnt);
}

static int stencil(const orc_grid* G, const int* cs, int cell, int* rs, int* re)
{
    int cz = cell / G->sheet, rem = cell % G->sheet, cy = rem / G->dx, cx = rem % G->dx;
    int k = 0;
    for (int dz = -1; dz <= 1; dz++) {
        int zz = cz + dz;
        if (zz < 0 || zz >= G->dz) continue;
        for (int dy = -1; dy <= 1; dy++) {
            int yy = cy + dy;
            if (yy < 0 || yy >= G->dy) continue;
            for (int dx = -1; dx <= 1; dx++) {
                int xx = cx + dx;
                if (xx < 0 || xx >= G->dx) continue;
                int c = zz * G->sheet + yy * G->dx + xx;
                rs[k] = cs[c]; re[k] = cs[c + 1]; k++;
            }
        }
    }
    return k;
}

/* cell of sorted particle i: cell_start is non-decreasing, find c with cs[c] <= i < cs[c+1] */
static int cell_of(const int* cs, int ncells, int i)
{
    int lo = 0, hi = ncells - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (cs[mid] <= i) lo = mid; else hi = mid - 1;
    }
    return lo;
}

/* Squared distance in the mode's precision: real3 dr = pos_j - pos_i; dr * dr
 * (MPS_CAL_expl_v1.0/typedef.h:51-54, left-to-right sum, no contraction). */
static inline double rr_mode(const orc_params* P, const double* x, const double* y,
                             const double* z, int i, int j)
{
    if (P->f32) {
        float dx = (float)x[j] - (float)x[i];
        float dy = (float)y[j] - (float)y[i];
        float dz = (float)z[j] - (float)z[i];
        float a = dx * dx, b = dy * dy, c = dz * dz;
        float s = a + b;
        s = s + c;
        return s;
    }
    double dx = x[j] - x[i], dy = y[j] - y[i], dz = z[j] - z[i];
    return dx * dx + dy * dy + dz * dz;
}

/* d_weight cutoff: w = 0 when sqrt(rr) >= re (MPS_CAL_expl_v1.0/MPS.cpp:301-306). */
static inline int inside(const orc_params* P, double rr, double re)
{
    if (P->f32) return sqrtf((float)rr) < (float)re;
    return sqrt(rr) < re;
}

static inline double weight(double re, double r) { return re / r - 1.0; }

/* Ordered within-cutoff neighbour lists (j != i), which = 0: R0, 1: RLAP.
 * Rows in sorted order, columns in the reference visiting order (ascending). */
int orc_neighbors(const orc_params* P, const orc_grid* G, int np, const double* x,
                  const double* y, const double* z, const int* cs, int which,
                  int* row_ptr, int* col, long cap)
{
    double th[6];
    orc_thresholds(P, th);
    double re = which ? th[1] : th[0];
    long k = 0;
    int rs[27], rend[27];
    row_ptr[0] = 0;
    for (int i = 0; i < np; i++) {
        int c = cell_of(cs, G->ncells, i);
        int m = stencil(G, cs, c, rs, rend);
        for (int s = 0; s < m; s++)
            for (int j = rs[s]; j < rend[s]; j++) {
                if (j == i) continue;
                if (inside(P, rr_mode(P, x, y, z, i, j), re)) {
                    if (k < cap) col[k] = j;
                    k++;
                }
            }
        row_ptr[i + 1] = (int)k;
    }
    return k <= cap ? 0 : -1;
}

/* ------------------------------------------------------------------ */
/* operators                                                             */
/* ------------------------------------------------------------------ */

/* MPS::calOnce (MPS_CAL_expl_v1.0/MPS.cpp:311-451): n_i over all particles,
 * n0 = max n_i (first maximum in index order), lambda = sum r^2 w / n0
 * around the argmax, and for every type-2 particle the nearest type-1
 * particle with |r|^2 < 2 R0^2 (first strict minimum in index order;
 * -1 when there is none, where the reference leaves the slot unset). */
void orc_calonce(const orc_params* P, const orc_grid* G, int np, const double* x,
                 const double* y, const double* z, const int* type, const int* cs,
                 double* n, double* n0_out, double* lambda_out, int* argmax_out, int* normal)
{
    double th[6];
    orc_thresholds(P, th);
    const double re = th[0];
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < np; i++) {
        int rs[27], rend[27];
        int m = stencil(G, cs, cell_of(cs, G->ncells, i), rs, rend);
        double acc = 0.0;
        for (int s = 0; s < m; s++)
            for (int j = rs[s]; j < rend[s]; j++) {
                if (j == i) continue;
                double rr = rr_mode(P, x, y, z, i, j);
                if (!inside(P, rr, re)) continue;
                double dx = x[j] - x[i], dy = y[j] - y[i], dz = z[j] - z[i];
                acc += weight(re, sqrt(dx * dx + dy * dy + dz * dz));
            }
        n[i] = acc;
    }
    double n0 = 0.0;
    int am = 0;
    for (int i = 0; i < np; i++)
        if (n[i] > n0) { n0 = n[i]; am = i; }
    double lam = 0.0;
    {
        int rs[27], rend[27];
        int m = stencil(G, cs, cell_of(cs, G->ncells, am), rs, rend);
        for (int s = 0; s < m; s++)
            for (int j = rs[s]; j < rend[s]; j++) {
                if (j == am) continue;
                double rr = rr_mode(P, x, y, z, am, j);
                if (!inside(P, rr, re)) continue;
                double dx = x[j] - x[am], dy = y[j] - y[am], dz = z[j] - z[am];
                double r = sqrt(dx * dx + dy * dy + dz * dz);
                lam += r * r * weight(re, r);
            }
    }
    *n0_out = n0;
    *lambda_out = lam / n0;
    *argmax_out = am;
    if (!normal) return;
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < np; i++) {
        if (type[i] != 2) { normal[i] = -1; continue; }
        int rs[27], rend[27];
        int m = stencil(G, cs, cell_of(cs, G->ncells, i), rs, rend);
        double best;
        if (P->f32) { float r0 = (float)th[0]; best = (float)(2.0f * r0) * r0; }
        else best = 2.0 * th[0] * th[0];
        int bj = -1;
        for (int s = 0; s < m; s++)
            for (int j = rs[s]; j < rend[s]; j++) {
                if (type[j] != 1) continue;
                double rr = rr_mode(P, x, y, z, i, j);
                if (rr < best) { best = rr; bj = j; }
            }
        normal[i] = bj;
    }
}

/* MPS_GPU::cal_n, CPU branch (MPS_CAL_expl_v1.0/MPS_GPU.cpp:843-885):
 * n_i = sum_{j != i} w(|r_ij|; R0) for type != 2; type-2 entries untouched. */
void orc_pnd(const orc_params* P, const orc_grid* G, int np, const double* x, const double* y,
             const double* z, const int* type, const int* cs, double* n)
{
    double th[6];
    orc_thresholds(P, th);
    const double re = th[0];
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < np; i++) {
        if (type[i] == 2) continue;
        int rs[27], rend[27];
        int m = stencil(G, cs, cell_of(cs, G->ncells, i), rs, rend);
        double acc = 0.0;
        for (int s = 0; s < m; s++)
            for (int j = rs[s]; j < rend[s]; j++) {
                if (j == i) continue;
                double rr = rr_mode(P, x, y, z, i, j);
                if (!inside(P, rr, re)) continue;
                double dx = x[j] - x[i], dy = y[j] - y[i], dz = z[j] - z[i];
                acc += weight(re, sqrt(dx * dx + dy * dy + dz * dz));
            }
        n[i] = acc;
    }
}

/* MPS_GPU_EXPL::calPres_expl, CPU branch (MPS_CAL_expl_v1.0/MPS_GPU_EXPL.cpp:199-238):
 * p = max(0, rho c^2 (n - n0) / n0) for type != 2, then dummy p = p[normal].
 * A dummy with no wall partner (normal < 0) gets p = 0. */
void orc_pres_expl(const orc_params* P, int np, const double* n, double n0, const int* type,
                   const int* normal, double* p)
{
    const double one_over_alpha = P->rho * P->cs * P->cs;
    for (int i = 0; i < np; i++)
        if (type[i] != 2) {
            double t = one_over_alpha * (n[i] - n0) / n0;
            p[i] = t > 0.0 ? t : 0.0;
        }
    for (int i = 0; i < np; i++)
        if (type[i] == 2) p[i] = normal[i] >= 0 ? p[normal[i]] : 0.0;
}

/* MPS_GPU::d3_LapVel (MPS_CAL_expl_v1.0/MPS_GPU.cpp:683-717; identical in
 * MPS_CAL_v4.0/MPS_CPU.cpp:490-525): coef sum_{j != i} w(|r|; RLAP)(u_j - u_i),
 * all types j. */
static void lapvel_one(const orc_params* P, const orc_grid* G, const double* x, const double* y,
                       const double* z, const double* u, const double* v, const double* w,
                       const int* cs, double re, double coef, int i, double out[3])
{
    int rs[27], rend[27];
    int m = stencil(G, cs, cell_of(cs, G->ncells, i), rs, rend);
    double lx = 0, ly = 0, lz = 0;
    for (int s = 0; s < m; s++)
        for (int j = rs[s]; j < rend[s]; j++) {
            if (j == i) continue;
            double rr = rr_mode(P, x, y, z, i, j);
            if (!inside(P, rr, re)) continue;
            double dx = x[j] - x[i], dy = y[j] - y[i], dz = z[j] - z[i];
            double wt = weight(re, sqrt(dx * dx + dy * dy + dz * dz));
            lx += wt * (u[j] - u[i]); ly += wt * (v[j] - v[i]); lz += wt * (w[j] - w[i]);
        }
    out[0] = coef * lx; out[1] = coef * ly; out[2] = coef * lz;
}

void orc_lapvel(const orc_params* P, const orc_grid* G, int np, const double* x, const double* y,
                const double* z, const double* u, const double* v, const double* w,
                const int* cs, double coef, double* lx, double* ly, double* lz)
{
    double th[6];
    orc_thresholds(P, th);
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < np; i++) {
        double o[3];
        lapvel_one(P, G, x, y, z, u, v, w, cs, th[1], coef, i, o);
        lx[i] = o[0]; ly[i] = o[1]; lz[i] = o[2];
    }
}

/* MPS_GPU_EXPL::calVisc_expl, CPU branch (MPS_CAL_expl_v1.0/MPS_GPU_EXPL.cpp:152-194),
 * two-pass: u* = u + dt (nu lap(u) + g), x* = x + dt u* for type 0; every
 * other particle is copied. coef = 2 d / (n0 lambda). */
void orc_visc(const orc_params* P, const orc_grid* G, int np, const double* x, const double* y,
              const double* z, const double* u, const double* v, const double* w,
              const int* type, const int* cs, double coef, double dt,
              double* xo, double* yo, double* zo, double* uo, double* vo, double* wo)
{
    double th[6];
    orc_thresholds(P, th);
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < np; i++) {
        xo[i] = x[i]; yo[i] = y[i]; zo[i] = z[i]; uo[i] = u[i]; vo[i] = v[i]; wo[i] = w[i];
        if (type[i] != 0) continue;
        double l[3];
        lapvel_one(P, G, x, y, z, u, v, w, cs, th[1], coef, i, l);
        double us = u[i] + dt * (P->nu * l[0] + P->g[0]);
        double vs = v[i] + dt * (P->nu * l[1] + P->g[1]);
        double ws = w[i] + dt * (P->nu * l[2] + P->g[2]);
        uo[i] = us; vo[i] = vs; wo[i] = ws;
        xo[i] = x[i] + dt * us; yo[i] = y[i] + dt * vs; zo[i] = z[i] + dt * ws;
    }
}

/* MPS_GPU::collision, CPU branch (MPS_CAL_expl_v1.0/MPS_GPU.cpp:892-986), two-pass:
 * for type-0 i and every j != i in the stencil, ds = |r_ij|,
 * v = 0.5 (u_j - u_i).r_ij / ds; when ds <= col_dis and v <= 0,
 * du_i += rate v r_ij / ds.  Then u += du, x += dt du for all i.
 * The accept/reject decision is taken in the mode's precision. */
int orc_collision(const orc_params* P, const orc_grid* G, int np, const double* x,
                  const double* y, const double* z, const double* u, const double* v,
                  const double* w, const int* type, const int* cs, double dt,
                  double* xo, double* yo, double* zo, double* uo, double* vo, double* wo)
{
    double th[6];
    orc_thresholds(P, th);
    const double cd = th[2];
    int ncol = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : ncol)
    for (int i = 0; i < np; i++) {
        double cx = 0, cy = 0, cz = 0;
        if (type[i] == 0) {
            int rs[27], rend[27];
            int m = stencil(G, cs, cell_of(cs, G->ncells, i), rs, rend);
            for (int s = 0; s < m; s++)
                for (int j = rs[s]; j < rend[s]; j++) {
                    if (j == i) continue;
                    int hit;
                    if (P->f32) {
                        float dx = (float)x[j] - (float)x[i], dy = (float)y[j] - (float)y[i],
                              dz = (float)z[j] - (float)z[i];
                        float du = (float)u[j] - (float)u[i], dv = (float)v[j] - (float)v[i],
                              dw = (float)w[j] - (float)w[i];
                        float rr = dx * dx + dy * dy;
                        rr = rr + dz * dz;
                        float ds = sqrtf(rr);
                        float ods = 1.0f / ds;
                        float hx = 0.5f * du, hy = 0.5f * dv, hz = 0.5f * dw;
                        float dot = hx * dx + hy * dy;
                        dot = dot + hz * dz;
                        float vabs = dot * ods;
                        hit = (ds <= (float)cd) && (vabs <= 0.0f);
                    } else {
                        double dx = x[j] - x[i], dy = y[j] - y[i], dz = z[j] - z[i];
                        double ds = sqrt(dx * dx + dy * dy + dz * dz);
                        double vabs = 0.5 * ((u[j] - u[i]) * dx + (v[j] - v[i]) * dy +
                                             (w[j] - w[i]) * dz) / ds;
                        hit = (ds <= cd) && (vabs <= 0.0);
                    }
                    if (!hit) continue;
                    double dx = x[j] - x[i], dy = y[j] - y[i], dz = z[j] - z[i];
                    double ds = sqrt(dx * dx + dy * dy + dz * dz);
                    double vabs = 0.5 * ((u[j] - u[i]) * dx + (v[j] - v[i]) * dy +
                                         (w[j] - w[i]) * dz) / ds;
                    double f = P->col_rate * vabs / ds;
                    cx += f * dx; cy += f * dy; cz += f * dz;
                    ncol++;
                }
        }
        uo[i] = u[i] + cx; vo[i] = v[i] + cy; wo[i] = w[i] + cz;
        xo[i] = x[i] + dt * cx; yo[i] = y[i] + dt * cy; zo[i] = z[i] + dt * cz;
    }
    return ncol;
}

/* MPS_GPU::d3_GradPres (MPS_CAL_expl_v1.0/MPS_GPU.cpp:571-638):
 * p_hat = min(p_i, p_j : type_j != 2, |r|^2 <= R0^2);
 * grad = (d / n0) sum_{j != i} (p_j - p_hat) / |r|^2 w(|r|; R0) r_ij.
 * gradp_skip_dummy drops type-2 j from the sum, as MPS_CPU::d3_GradPress
 * does (MPS_CAL_v4.0/MPS_CPU.cpp:384-445). */
static void gradp_one(const orc_params* P, const orc_grid* G, const double* x, const double* y,
                      const double* z, const double* p, const int* type, const int* cs,
                      double re, double re2, double n0, int i, double out[3])
{
    int rs[27], rend[27];
    int m = stencil(G, cs, cell_of(cs, G->ncells, i), rs, rend);
    double ph = p[i];
    for (int s = 0; s < m; s++)
        for (int j = rs[s]; j < rend[s]; j++) {
            if (type[j] == 2) continue;
            double rr = rr_mode(P, x, y, z, i, j);
            int near = P->f32 ? ((float)rr <= (float)re2) : (rr <= re2);
            if (near && p[j] < ph) ph = p[j];
        }
    double gx = 0, gy = 0, gz = 0;
    for (int s = 0; s < m; s++)
        for (int j = rs[s]; j < rend[s]; j++) {
            if (j == i) continue;
            if (P->gradp_skip_dummy && type[j] == 2) continue;
            double rrm = rr_mode(P, x, y, z, i, j);
            if (!inside(P, rrm, re)) continue;
            double dx = x[j] - x[i], dy = y[j] - y[i], dz = z[j] - z[i];
            double rr = dx * dx + dy * dy + dz * dz;
            double f = (p[j] - ph) / rr * weight(re, sqrt(rr));
            gx += f * dx; gy += f * dy; gz += f * dz;
        }
    double s = (double)P->dim / n0;
    out[0] = s * gx; out[1] = s * gy; out[2] = s * gz;
}

void orc_gradp_vec(const orc_params* P, const orc_grid* G, int np, const double* x,
                   const double* y, const double* z, const double* p, const int* type,
                   const int* cs, double n0, double* gx, double* gy, double* gz)
{
    double th[6];
    orc_thresholds(P, th);
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < np; i++) {
        double o[3];
        gradp_one(P, G, x, y, z, p, type, cs, th[0], th[3], n0, i, o);
        gx[i] = o[0]; gy[i] = o[1]; gz[i] = o[2];
    }
}

/* MPS_GPU::calDash, CPU branch (MPS_CAL_expl_v1.0/MPS_GPU.cpp:782-825), two-pass:
 * type 0: du = -dt / rho grad p, u += du, x += dt du. */
void orc_gradp(const orc_params* P, const orc_grid* G, int np, const double* x, const double* y,
               const double* z, const double* u, const double* v, const double* w,
               const double* p, const int* type, const int* cs, double n0, double dt,
               double* xo, double* yo, double* zo, double* uo, double* vo, double* wo)
{
    double th[6];
    orc_thresholds(P, th);
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < np; i++) {
        xo[i] = x[i]; yo[i] = y[i]; zo[i] = z[i]; uo[i] = u[i]; vo[i] = v[i]; wo[i] = w[i];
        if (type[i] != 0) continue;
        double g[3];
        gradp_one(P, G, x, y, z, p, type, cs, th[0], th[3], n0, i, g);
        double s = -dt / P->rho;
        double du = s * g[0], dv = s * g[1], dw = s * g[2];
        uo[i] = u[i] + du; vo[i] = v[i] + dv; wo[i] = w[i] + dw;
        xo[i] = x[i] + dt * du; yo[i] = y[i] + dt * dv; zo[i] = z[i] + dt * dw;
    }
}

/* MPS_GPU::update_dt (MPS_CAL_expl_v1.0/MPS_GPU.cpp:761-777):
 * dt = min(dt_max, min_i CFL dp / |u_i|); too small when dt < dt_min. */
double orc_update_dt(const orc_params* P, int np, const double* u, const double* v,
                     const double* w, int* too_small)
{
    double th[6];
    orc_thresholds(P, th);
    double umax2 = 0.0;
    for (int i = 0; i < np; i++) {
        double a = u[i] * u[i] + v[i] * v[i] + w[i] * w[i];
        if (a > umax2) umax2 = a;
    }
    double dt = P->dt_max;
    if (umax2 > 0.0) {
        double t = P->cfl * P->dp / sqrt(umax2);
        if (t < dt) dt = t;
    }
    if (too_small) *too_small = dt < th[4];
    return dt;
}

/* MPS_CPU::buildPoisson (MPS_CAL_v4.0/MPS_CPU.cpp:619-710) in CSR form:
 * bc_i = n_i < beta n0; rows with bc_i or type 2 are identity with b = 0;
 * otherwise A_ij = w(|r|; RLAP) for type_j != 2, j not bc, and
 * A_ii = -sum_{j != i, type_j != 2} w_ij - rho lambda alpha n0 / (2 d dt^2);
 * b_i = lambda rho / (2 d dt^2) (n0 - n_i); x0 = p_prev (0 on bc rows).
 * Zero entries of the dense matrix are not stored; columns ascend.
 * Returns nnz, or -nnz when cap is too small. */
long orc_build_poisson(const orc_params* P, const orc_grid* G, int np, const double* x,
                       const double* y, const double* z, const int* type, const int* cs,
                       const double* n, const double* p_prev, double n0, double lambda,
                       double dt, double alpha, int* row_ptr, int* col, double* val,
                       double* b, double* x0, unsigned char* bc, long cap)
{
    double th[6];
    orc_thresholds(P, th);
    const double re = th[1];
    const double d = (double)P->dim;
    const double aii = P->rho * lambda * alpha * n0 / (2.0 * d * dt * dt);
    const double btmp = lambda * P->rho * (0.5 * (1.0 / d) * (1.0 / dt) * (1.0 / dt));
    for (int i = 0; i < np; i++) {
        b[i] = btmp * (n0 - n[i]);
        x0[i] = p_prev ? p_prev[i] : 0.0;
        bc[i] = n[i] < P->beta * n0;
        if (bc[i]) x0[i] = 0.0;
    }
    long k = 0;
    int rs[27], rend[27];
    row_ptr[0] = 0;
    for (int i = 0; i < np; i++) {
        if (bc[i] || type[i] == 2) {
            b[i] = 0.0;
            if (k < cap) { col[k] = i; val[k] = 1.0; }
            k++;
            row_ptr[i + 1] = (int)k;
            continue;
        }
        int m = stencil(G, cs, cell_of(cs, G->ncells, i), rs, rend);
        double diag = 0.0;
        long kd = -1;
        for (int s = 0; s < m; s++)
            for (int j = rs[s]; j < rend[s]; j++) {
                if (j == i) {
                    kd = k;
                    if (k < cap) { col[k] = i; val[k] = 0.0; }
                    k++;
                    continue;
                }
                if (type[j] == 2) continue;
                double rrm = rr_mode(P, x, y, z, i, j);
                if (!inside(P, rrm, re)) continue;
                double dx = x[j] - x[i], dy = y[j] - y[i], dz = z[j] - z[i];
                double wt = weight(re, sqrt(dx * dx + dy * dy + dz * dz));
                diag -= wt;
                if (bc[j]) continue;
                if (k < cap) { col[k] = j; val[k] = wt; }
                k++;
            }
        if (kd >= 0 && kd < cap) val[kd] = diag - aii;
        row_ptr[i + 1] = (int)k;
    }
    return k <= cap ? k : -k;
}

void orc_spmv(int np, const int* rp, const int* col, const double* val, const double* x, double* y)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < np; i++) {
        double s = 0.0;
        for (int k = rp[i]; k < rp[i + 1]; k++) s += val[k] * x[col[k]];
        y[i] = s;
    }
}

static double dot(int n, const double* a, const double* b)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* MPS_CPU::DoCG (MPS_CAL_v4.0/MPS_CPU.cpp:829-888): textbook CG from x.
 * Stops when r.r <= tol (relative = 0, the reference's EPS_BY_EPS test,
 * common.h:45-48) or when r.r <= tol^2 b.b (relative = 1).  Returns the
 * iteration count; -1 when maxit is reached first. */
int orc_cg(int np, const int* rp, const int* col, const double* val, const double* b,
           double* x, double tol, int relative, int maxit, double* rr_out)
{
    double* r = (double*)malloc(sizeof(double) * (size_t)np);
    double* p = (double*)malloc(sizeof(double) * (size_t)np);
    double* ap = (double*)malloc(sizeof(double) * (size_t)np);
    orc_spmv(np, rp, col, val, x, ap);
    for (int i = 0; i < np; i++) { r[i] = b[i] - ap[i]; p[i] = r[i]; }
    double rr = dot(np, r, r);
    double stop = relative ? tol * tol * dot(np, b, b) : tol;
    int it = 0;
    while (rr > stop) {
        if (it >= maxit) { it = -1; break; }
        orc_spmv(np, rp, col, val, p, ap);
        double a = rr / dot(np, ap, p);
        for (int i = 0; i < np; i++) { x[i] += a * p[i]; r[i] -= a * ap[i]; }
        double rn = dot(np, r, r);
        double bt = rn / rr;
        for (int i = 0; i < np; i++) p[i] = r[i] + bt * p[i];
        rr = rn;
        it++;
    }
    if (rr_out) *rr_out = rr;
    free(r); free(p); free(ap);
    return it;
}

/* ------------------------------------------------------------------ */
/* whole-step simulator                                                  */
/* ------------------------------------------------------------------ */

struct orc_sim {
    orc_params P;
    orc_grid G;
    int nt, np;                        /* allocated, active */
    double *x, *y, *z, *u, *v, *w, *p, *n;
    double *x2, *y2, *z2, *u2, *v2, *w2, *tmp;
    int *type, *id, *normal, *key, *perm, *index, *cs, *itmp;
    double n0, lambda, coef, dt, time;
    /* prescribed motion: flags and reference coordinate by particle id */
    unsigned char* mov;
    double* xref;
    int mov_axis;
    double mov_amp, mov_omega;
};

static void permute_d(int nt, const int* perm, double* a, double* tmp)
{
    for (int k = 0; k < nt; k++) tmp[k] = a[perm[k]];
    memcpy(a, tmp, sizeof(double) * (size_t)nt);
}

static void permute_i(int nt, const int* perm, int* a, int* tmp)
{
    for (int k = 0; k < nt; k++) tmp[k] = a[perm[k]];
    memcpy(a, tmp, sizeof(int) * (size_t)nt);
}

/* MPS_GPU::divideCell + sort_* + sort_normal (MPS_CAL_expl_v1.0/MPS_GPU.cpp:248-373). */
static void sim_divide(orc_sim* S)
{
    orc_hash(&S->P, &S->G, S->nt, S->x, S->y, S->z, S->type, S->key);
    for (int i = 0; i < S->nt; i++)
        if (S->key[i] == S->G.ncells && S->type[i] != 3) {
            S->type[i] = 3;
            S->u[i] = S->v[i] = S->w[i] = 0.0;
        }
    orc_sort(S->nt, S->G.ncells, S->key, S->perm, S->index, S->cs);
    double* a[] = {S->x, S->y, S->z, S->u, S->v, S->w, S->p, S->n};
    for (int q = 0; q < 8; q++) permute_d(S->nt, S->perm, a[q], S->tmp);
    permute_i(S->nt, S->perm, S->type, S->itmp);
    permute_i(S->nt, S->perm, S->id, S->itmp);
    for (int k = 0; k < S->nt; k++) {
        int o = S->normal[S->perm[k]];
        S->itmp[k] = o >= 0 ? S->index[o] : -1;
    }
    memcpy(S->normal, S->itmp, sizeof(int) * (size_t)S->nt);
    S->np = S->cs[S->G.ncells];
}

orc_sim* orc_sim_create(const orc_params* P, int np, int nb2, int nb1, const double* x,
                        const double* y, const double* z, const double* u, const double* v,
                        const double* w, const double* p)
{
    orc_sim* S = (orc_sim*)calloc(1, sizeof(orc_sim));
    S->P = *P;
    S->nt = np;
    size_t nd = sizeof(double) * (size_t)np, ni = sizeof(int) * (size_t)np;
    double** dd[] = {&S->x, &S->y, &S->z, &S->u, &S->v, &S->w, &S->p, &S->n,
                     &S->x2, &S->y2, &S->z2, &S->u2, &S->v2, &S->w2, &S->tmp};
    for (int q = 0; q < 15; q++) *dd[q] = (double*)calloc((size_t)np, sizeof(double));
    int** ii[] = {&S->type, &S->id, &S->normal, &S->key, &S->perm, &S->index, &S->itmp};
    for (int q = 0; q < 7; q++) *ii[q] = (int*)calloc((size_t)np, sizeof(int));
    memcpy(S->x, x, nd); memcpy(S->y, y, nd); memcpy(S->z, z, nd);
    memcpy(S->u, u, nd); memcpy(S->v, v, nd); memcpy(S->w, w, nd); memcpy(S->p, p, nd);
    (void)ni;
    /* MPS::createGeo (MPS_CAL_expl_v1.0/MPS.cpp:252-268): id = i, types by index range */
    for (int i = 0; i < np; i++) {
        S->id[i] = i;
        S->type[i] = i < nb2 ? 2 : (i < nb2 + nb1 ? 1 : 0);
        S->normal[i] = -1;
    }
    if (orc_grid_init(P, np, x, y, z, &S->G) != 0) { orc_sim_destroy(S); return NULL; }
    S->cs = (int*)calloc((size_t)S->G.ncells + 2, sizeof(int));
    /* MPS_GPU::Initial order (MPS_CAL_expl_v1.0/MPS_GPU.cpp:100-106) */
    sim_divide(S);
    int am;
    orc_calonce(&S->P, &S->G, S->np, S->x, S->y, S->z, S->type, S->cs, S->n, &S->n0,
                &S->lambda, &am, S->normal);
    S->coef = 2.0 * P->dim / (S->n0 * S->lambda);
    S->dt = P->dt_max;
    S->time = 0.0;
    return S;
}

void orc_sim_destroy(orc_sim* S)
{
    if (!S) return;
    double* dd[] = {S->x, S->y, S->z, S->u, S->v, S->w, S->p, S->n,
                    S->x2, S->y2, S->z2, S->u2, S->v2, S->w2, S->tmp};
    for (int q = 0; q < 15; q++) free(dd[q]);
    int* ii[] = {S->type, S->id, S->normal, S->key, S->perm, S->index, S->itmp, S->cs};
    for (int q = 0; q < 8; q++) free(ii[q]);
    free(S->mov);
    free(S->xref);
    free(S);
}

/* The operators write the np particles inside cells; the excluded tail
 * [np, nt) is carried into the new buffer so that an excluded particle keeps
 * the position it left the grid at (the reference overwrites slot np-1
 * instead, MPS_GPU.cpp:311-314). */
static void swap_state(orc_sim* S)
{
    const size_t tail = sizeof(double) * (size_t)(S->nt - S->np);
    if (tail) {
        memcpy(S->x2 + S->np, S->x + S->np, tail);
        memcpy(S->y2 + S->np, S->y + S->np, tail);
        memcpy(S->z2 + S->np, S->z + S->np, tail);
        memcpy(S->u2 + S->np, S->u + S->np, tail);
        memcpy(S->v2 + S->np, S->v + S->np, tail);
        memcpy(S->w2 + S->np, S->w + S->np, tail);
    }
    double* t;
    t = S->x; S->x = S->x2; S->x2 = t;
    t = S->y; S->y = S->y2; S->y2 = t;
    t = S->z; S->z = S->z2; S->z2 = t;
    t = S->u; S->u = S->u2; S->u2 = t;
    t = S->v; S->v = S->v2; S->v2 = t;
    t = S->w; S->w = S->w2; S->w2 = t;
}

static double st_round(const orc_sim* S, double v) { return S->P.f32 ? (double)(float)v : v; }

static double* axis_of(orc_sim* S, int a, int vel)
{
    if (vel) return a == 0 ? S->u : (a == 1 ? S->v : S->w);
    return a == 0 ? S->x : (a == 1 ? S->y : S->z);
}

/* Prescribed sinusoidal motion (the reference's MOTION::doMotion is an empty
 * stub, MPS_CAL_/MOTION.cpp:11-16): flagged particles (by id) follow
 * x_axis(t) = xref + A sin(w t), velocity A w cos(w t) along the axis, zero
 * across it; xref is taken so that the law continues from the current state. */
void orc_sim_set_motion(orc_sim* S, const unsigned char* moving_by_id, int axis, double amp, double freq)
{
    free(S->mov); free(S->xref);
    S->mov = NULL; S->xref = NULL; S->mov_amp = 0.0;
    if (!moving_by_id || amp == 0.0) return;
    S->mov = (unsigned char*)malloc((size_t)S->nt);
    S->xref = (double*)calloc((size_t)S->nt, sizeof(double));
    memcpy(S->mov, moving_by_id, (size_t)S->nt);
    S->mov_axis = axis; S->mov_amp = amp; S->mov_omega = 2.0 * 3.14159265358979323846 * freq;
    double* xa = axis_of(S, axis, 0);
    const double va = st_round(S, amp * S->mov_omega * cos(S->mov_omega * S->time));
    for (int i = 0; i < S->nt; i++) {
        const int o = S->id[i];
        if (!S->mov[o]) continue;
        S->xref[o] = xa[i] - amp * sin(S->mov_omega * S->time);
        for (int a = 0; a < 3; a++) axis_of(S, a, 1)[i] = a == axis ? va : 0.0;
    }
}

static void sim_motion(orc_sim* S, double t)
{
    if (!S->mov || S->mov_amp == 0.0) return;
    double* xa = axis_of(S, S->mov_axis, 0);
    const double pos_off = S->mov_amp * sin(S->mov_omega * t);
    const double va = st_round(S, S->mov_amp * S->mov_omega * cos(S->mov_omega * t));
    for (int i = 0; i < S->nt; i++) {
        const int o = S->id[i];
        if (!S->mov[o]) continue;
        xa[i] = st_round(S, S->xref[o] + pos_off);
        for (int a = 0; a < 3; a++) axis_of(S, a, 1)[i] = a == S->mov_axis ? va : 0.0;
    }
}

/* One explicit step.  rebin_every_move = 0 follows MPS_CAL_/MPS_GPU.cpp:611-714
 * (one re-binning at the start of the step); 1 follows
 * MPS_CAL_expl_v1.0/MPS_GPU_EXPL.cpp:85-147 (re-binning after every position
 * update).  dt is taken from the velocities at the start of the step
 * (MPS_GPU_EXPL.cpp:95).  Returns 1 when dt fell below dt_min. */
int orc_sim_step(orc_sim* S, int nsteps, int rebin_every_move)
{
    for (int s = 0; s < nsteps; s++) {
        int small = 0;
        S->dt = orc_update_dt(&S->P, S->np, S->u, S->v, S->w, &small);
        if (small) return 1;
        sim_motion(S, S->time + S->dt);
        if (!rebin_every_move) sim_divide(S);
        orc_visc(&S->P, &S->G, S->np, S->x, S->y, S->z, S->u, S->v, S->w, S->type, S->cs,
                 S->coef, S->dt, S->x2, S->y2, S->z2, S->u2, S->v2, S->w2);
        swap_state(S);
        if (rebin_every_move) sim_divide(S);
        orc_collision(&S->P, &S->G, S->np, S->x, S->y, S->z, S->u, S->v, S->w, S->type, S->cs,
                      S->dt, S->x2, S->y2, S->z2, S->u2, S->v2, S->w2);
        swap_state(S);
        if (rebin_every_move) sim_divide(S);
        orc_pnd(&S->P, &S->G, S->np, S->x, S->y, S->z, S->type, S->cs, S->n);
        orc_pres_expl(&S->P, S->np, S->n, S->n0, S->type, S->normal, S->p);
        orc_gradp(&S->P, &S->G, S->np, S->x, S->y, S->z, S->u, S->v, S->w, S->p, S->type, S->cs,
                  S->n0, S->dt, S->x2, S->y2, S->z2, S->u2, S->v2, S->w2);
        swap_state(S);
        if (rebin_every_move) sim_divide(S);
        S->time += S->dt;
    }
    return 0;
}

/* MPS_CPU_EXPL::step (MPS_CAL_v4.0/MPS_CPU_EXPL.cpp:79-135) with two-pass
 * operators, in the order of the device's semi-implicit step: dt, [re-bin],
 * viscosity + g, collision, n*, buildPoisson (MPS_CPU.cpp:619-710), FP64 CG
 * from the previous pressure to |b - A x| <= tol |b|, pressCorr p = max(x, 0)
 * (MPS_CPU.cpp:734-743), pressure gradient without type-2 neighbours
 * (MPS_CPU.cpp:384-445).  Returns 0, 1 when dt fell below dt_min, 2 when CG
 * did not converge.  iters (may be NULL) receives the summed CG iterations. */
int orc_sim_step_semi(orc_sim* S, int nsteps, int rebin_every_move, double alpha, double tol, int* iters)
{
    const int n = S->nt;
    int* rp = (int*)malloc(sizeof(int) * ((size_t)n + 1));
    double* b = (double*)malloc(sizeof(double) * (size_t)n);
    double* x = (double*)malloc(sizeof(double) * (size_t)n);
    unsigned char* bc = (unsigned char*)malloc((size_t)n);
    int* col = NULL;
    double* val = NULL;
    long cap = 0;
    int total = 0, rc = 0;
    for (int s = 0; s < nsteps && !rc; s++) {
        int small = 0;
        S->dt = orc_update_dt(&S->P, S->np, S->u, S->v, S->w, &small);
        if (small) { rc = 1; break; }
        sim_motion(S, S->time + S->dt);
        if (!rebin_every_move) sim_divide(S);
        orc_visc(&S->P, &S->G, S->np, S->x, S->y, S->z, S->u, S->v, S->w, S->type, S->cs,
                 S->coef, S->dt, S->x2, S->y2, S->z2, S->u2, S->v2, S->w2);
        swap_state(S);
        if (rebin_every_move) sim_divide(S);
        orc_collision(&S->P, &S->G, S->np, S->x, S->y, S->z, S->u, S->v, S->w, S->type, S->cs,
                      S->dt, S->x2, S->y2, S->z2, S->u2, S->v2, S->w2);
        swap_state(S);
        if (rebin_every_move) sim_divide(S);
        orc_pnd(&S->P, &S->G, S->np, S->x, S->y, S->z, S->type, S->cs, S->n);
        long nnz;
        for (;;) {
            nnz = orc_build_poisson(&S->P, &S->G, S->np, S->x, S->y, S->z, S->type, S->cs, S->n, S->p,
                                    S->n0, S->lambda, S->dt, alpha, rp, col, val, b, x, bc, cap);
            if (nnz >= 0 && nnz <= cap) break;
            cap = (nnz < 0 ? -nnz : nnz) + 1024;
            free(col); free(val);
            col = (int*)malloc(sizeof(int) * (size_t)cap);
            val = (double*)malloc(sizeof(double) * (size_t)cap);
        }
        const int it = orc_cg(S->np, rp, col, val, b, x, tol, 1, 100000, NULL);
        if (it < 0) { rc = 2; break; }
        total += it;
        for (int i = 0; i < S->np; i++) S->p[i] = x[i] > 0.0 ? x[i] : 0.0;
        if (S->P.f32)
            for (int i = 0; i < S->np; i++) S->p[i] = (double)(float)S->p[i];
        const int keep = S->P.gradp_skip_dummy;
        S->P.gradp_skip_dummy = 1;
        orc_gradp(&S->P, &S->G, S->np, S->x, S->y, S->z, S->u, S->v, S->w, S->p, S->type, S->cs,
                  S->n0, S->dt, S->x2, S->y2, S->z2, S->u2, S->v2, S->w2);
        S->P.gradp_skip_dummy = keep;
        swap_state(S);
        if (rebin_every_move) sim_divide(S);
        S->time += S->dt;
    }
    if (iters) *iters = total;
    free(rp); free(b); free(x); free(bc); free(col); free(val);
    return rc;
}

int orc_sim_np(const orc_sim* S) { return S->np; }

void orc_sim_get(const orc_sim* S, double* x, double* y, double* z, double* u, double* v,
                 double* w, double* p, double* n, int* type, int* id)
{
    size_t nd = sizeof(double) * (size_t)S->nt, ni = sizeof(int) * (size_t)S->nt;
    if (x) memcpy(x, S->x, nd);
    if (y) memcpy(y, S->y, nd);
    if (z) memcpy(z, S->z, nd);
    if (u) memcpy(u, S->u, nd);
    if (v) memcpy(v, S->v, nd);
    if (w) memcpy(w, S->w, nd);
    if (p) memcpy(p, S->p, nd);
    if (n) memcpy(n, S->n, nd);
    if (type) memcpy(type, S->type, ni);
    if (id) memcpy(id, S->id, ni);
}

void orc_sim_consts(const orc_sim* S, double* out)
{
    out[0] = S->n0; out[1] = S->lambda; out[2] = S->dt; out[3] = S->time; out[4] = S->coef;
    out[5] = S->G.ncells;
}

void orc_sim_grid(const orc_sim* S, orc_grid* G) { *G = S->G; }
