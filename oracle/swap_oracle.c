/*
 * swap_oracle.c -- CPU restatement of SWAP's hybrid-PIC time step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never the
 * product.  Each routine cites the reference routine (file:line) whose
 * algorithm it restates.  Loops are written with the reference's 1-based
 * Fortran indices through the F3/F4 macros so every index can be checked
 * against the Fortran text.
 *
 * Compiled twice: OR_REAL=double (FP64 reference arithmetic) and OR_REAL=float (the
 * reference's own "-real-size 32" width, makefile:1).  Every literal goes
 * through R() so the float build really computes in float.
 */
#include "swap_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define R(x) ((real)(x))
/* parallel loop over particles or k-planes (every such loop writes only its
   own particle or plane) */
#define OMP_FOR _Pragma("omp parallel for schedule(static) num_threads(h->nthreads) if (h->nthreads > 1)")

struct orc {
  int nx, ny, nz;
  real dx, dy, delz, dt, dtsub, ntf, alpha, beta;
  int ntsub;
  real *qx, *qy, *qz, *dx_grid, *dy_grid, *dz_grid;
  real *dx_cell, *dy_cell, *dz_cell, *xrat, *yrat, *zrat;
  /* module scratch of global.f:59 (a, c, ct) */
  real *a, *c, *ct;
  /* open x (build-defined inflow/outflow; the reference is periodic) and
     its inflow (or_set_inflow): seed, upstream ions per cell, drift,
     thermal speed, beta_p, the id of the next ion; step counts or_step
     calls (it keys the inflow draws) */
  int ox;
  uint64_t in_seed;
  double in_ppc, in_vsw, in_vth, in_bp;
  uint32_t next_id;
  int64_t step;
  /* OpenMP threads of the particle and grid loops (or_set_threads; 1: the
     serial restatement).  Deposits then sum per-thread grids in thread
     order, so a given thread count is deterministic; counts differ in the
     rounding of the deposit sums only. */
  int nthreads;
};

/* Fortran (i,j,k) / (i,j,k,m), 1-based, column-major */
#define F3(i, j, k) ((size_t)((i)-1) + (size_t)h->nx * ((size_t)((j)-1) + (size_t)h->ny * (size_t)((k)-1)))
#define NN ((size_t)h->nx * h->ny * h->nz)
#define F4(i, j, k, m) (F3(i, j, k) + NN * (size_t)((m)-1))
/* 1-D coordinate arrays, 1-based */
#define QX(i) h->qx[(i)-1]
#define QY(j) h->qy[(j)-1]
#define QZ(k) h->qz[(k)-1]

int or_real_size(void) { return (int)sizeof(real); }

/* grd7, initial.f:147-301, with stretch factors xsf = zsf = 0 (initial.f:173,
   209), i.e. a uniform grid built as an FP cumulative sum (initial.f:193-196,
   233-237).  The reference's stretch loops also write dx_grid outside 1..nx
   when nx < 41 (initial.f:175-191); with xsf = 0 every value they write is dx,
   so the in-bounds result is the same as filling dx_grid with dx. */
static void grd7(orc *h) {
  int nx = h->nx, ny = h->ny, nz = h->nz;
  for (int j = 1; j <= ny; j++) {
    QY(j) = (real)j * h->dy; /* initial.f:159-162 */
    h->dy_grid[j - 1] = h->dy;
  }
  for (int i = 1; i <= nx; i++) h->dx_grid[i - 1] = h->dx;
  QX(1) = h->dx;
  for (int i = 2; i <= nx; i++) QX(i) = QX(i - 1) + h->dx_grid[i - 1];
  for (int i = 1; i <= nx - 1; i++) h->dx_grid[i - 1] = QX(i + 1) - QX(i);
  h->dx_grid[nx - 1] = h->dx_grid[nx - 2];
  for (int k = 1; k <= nz; k++) h->dz_grid[k - 1] = h->delz;
  QZ(1) = h->delz;
  for (int k = 2; k <= nz; k++) QZ(k) = QZ(k - 1) + h->dz_grid[k - 1];
  for (int k = 1; k <= nz - 1; k++) h->dz_grid[k - 1] = QZ(k + 1) - QZ(k);
  h->dz_grid[nz - 1] = h->dz_grid[nz - 2];
  /* cell sizes and interpolation ratios, initial.f:240-282 */
  h->dz_cell[0] = h->dz_grid[0];
  h->dz_cell[nz - 1] = h->dz_grid[nz - 1];
  h->zrat[0] = R(0.5);
  h->zrat[nz - 1] = R(0.5);
  for (int k = 2; k <= nz - 1; k++) {
    real zp = (QZ(k + 1) + QZ(k)) / R(2.0), zm = (QZ(k) + QZ(k - 1)) / R(2.0);
    h->dz_cell[k - 1] = zp - zm;
    h->zrat[k - 1] = (QZ(k) - zm) / (zp - zm);
  }
  h->dx_cell[0] = h->dx_grid[0];
  h->dx_cell[nx - 1] = h->dx_grid[nx - 1];
  h->xrat[0] = R(0.5);
  h->xrat[nx - 1] = R(0.5);
  for (int i = 2; i <= nx - 1; i++) {
    real xp = (QX(i + 1) + QX(i)) / R(2.0), xm = (QX(i) + QX(i - 1)) / R(2.0);
    h->dx_cell[i - 1] = xp - xm;
    h->xrat[i - 1] = (QX(i) - xm) / (xp - xm);
  }
  h->dy_cell[0] = h->dy_grid[0];
  h->dy_cell[ny - 1] = h->dy_grid[ny - 1];
  h->yrat[0] = R(0.5);
  h->yrat[ny - 1] = R(0.5);
  for (int j = 2; j <= ny - 1; j++) {
    real yp = (QY(j + 1) + QY(j)) / R(2.0), ym = (QY(j) + QY(j - 1)) / R(2.0);
    h->dy_cell[j - 1] = yp - ym;
    h->yrat[j - 1] = (QY(j) - ym) / (yp - ym);
  }
}

orc *or_create(int nx, int ny, int nz, double dx, double dy, double delz,
               double dt, int ntsub, double alpha, double beta) {
  orc *h = (orc *)calloc(1, sizeof(orc));
  h->nx = nx; h->ny = ny; h->nz = nz;
  h->dx = (real)dx; h->dy = (real)dy; h->delz = (real)delz;
  h->dt = (real)dt; h->ntsub = ntsub;
  h->dtsub = h->dt / (real)ntsub; /* inputs.f:145 */
  h->ntf = (real)ntsub;
  h->alpha = (real)alpha; h->beta = (real)beta;
  h->nthreads = 1;
  real **c1[] = {&h->qx, &h->qy, &h->qz, &h->dx_grid, &h->dy_grid, &h->dz_grid,
                 &h->dx_cell, &h->dy_cell, &h->dz_cell, &h->xrat, &h->yrat, &h->zrat};
  int n1[] = {nx, ny, nz, nx, ny, nz, nx, ny, nz, nx, ny, nz};
  for (int q = 0; q < 12; q++) *c1[q] = (real *)calloc((size_t)n1[q], sizeof(real));
  size_t n4 = (size_t)nx * ny * nz * 3;
  h->a = (real *)calloc(n4, sizeof(real));
  h->c = (real *)calloc(n4, sizeof(real));
  h->ct = (real *)calloc(n4, sizeof(real));
  grd7(h);
  return h;
}

void or_destroy(orc *h) {
  if (!h) return;
  real *p[] = {h->qx, h->qy, h->qz, h->dx_grid, h->dy_grid, h->dz_grid, h->dx_cell,
               h->dy_cell, h->dz_cell, h->xrat, h->yrat, h->zrat, h->a, h->c, h->ct};
  for (int q = 0; q < 15; q++) free(p[q]);
  free(h);
}

void or_get_coord(const orc *h, int which, real *out) {
  const real *src[] = {h->qx, h->qy, h->qz, h->dx_grid, h->dy_grid, h->dz_grid,
                       h->dx_cell, h->dy_cell, h->dz_cell, h->xrat, h->yrat, h->zrat};
  int n1[] = {h->nx, h->ny, h->nz, h->nx, h->ny, h->nz, h->nx, h->ny, h->nz, h->nx, h->ny, h->nz};
  if (which < 0 || which > 11) return;
  memcpy(out, src[which], sizeof(real) * (size_t)n1[which]);
}

void or_set_dtsub(orc *h, double dtsub, double ntf) { h->dtsub = (real)dtsub; h->ntf = (real)ntf; }
void or_set_open_x(orc *h, int on) { h->ox = on ? 1 : 0; }
void or_set_threads(orc *h, int n) { h->nthreads = n > 1 ? n : 1; }
void or_get_dtsub(const orc *h, double *dtsub, double *ntf) { *dtsub = h->dtsub; *ntf = h->ntf; }

/* initparameters, inputs.f:125-172 (constants inputs.f:19-67) */
void or_initparameters(double b0_init, int ion_amu, double nf_init, double dt_frac,
                       double dx_frac, double nu_init_frac, int ntsub, double *out) {
  const real q = R(1.6e-19), pi = R(3.14159);
  const real mu0 = pi * R(4.0e-7);
  real mion = (real)ion_amu * R(1.67e-27);
  real omega_p = q * (real)b0_init / mion;
  real lambda_i = (R(3e8) / (real)sqrt((double)(((real)nf_init / R(1e9)) * q * q /
                                                (R(8.85e-12) * mion)))) / R(1e3);
  real dx = lambda_i * (real)dx_frac;
  real dt = (real)dt_frac * mion / (q * (real)b0_init);
  real dtsub = dt / (real)ntsub;
  real nu_init = (real)nu_init_frac * omega_p;
  real alpha = (mu0 / R(1e3)) * q * (q / mion);
  double o[8] = {mion, omega_p, lambda_i, dx, dt, dtsub, nu_init, alpha};
  memcpy(out, o, sizeof(o));
}

/* grd6_setup, initial.f:36-63: uniform b0 = (0, 0, b0_init*q/mO) */
void or_grd6_b0(const orc *h, double b0_init, double mion, real *b0) {
  real eoverm = R(1.6e-19) / (real)mion;
  real bz = (real)b0_init * eoverm;
  for (size_t n = 0; n < NN; n++) {
    b0[n] = 0; b0[n + NN] = 0; b0[n + 2 * NN] = bz;
  }
}

/* periodic, boundary.f:9-52: x, then y, then z ghost copies.  An open x
   (build-defined) copies the nearest interior layer instead: ghost 1 takes
   layer 2, ghost nx takes layer nx-1 (zero gradient). */
void or_periodic(const orc *h, real *b) {
  int nx = h->nx, ny = h->ny, nz = h->nz;
  int sx1 = h->ox ? 2 : nx - 1, sxn = h->ox ? nx - 1 : 2;
  for (int m = 1; m <= 3; m++) {
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++) {
        b[F4(1, j, k, m)] = b[F4(sx1, j, k, m)];
        b[F4(nx, j, k, m)] = b[F4(sxn, j, k, m)];
      }
  }
  for (int m = 1; m <= 3; m++)
    for (int k = 1; k <= nz; k++)
      for (int i = 1; i <= nx; i++) {
        b[F4(i, 1, k, m)] = b[F4(i, ny - 1, k, m)];
        b[F4(i, ny, k, m)] = b[F4(i, 2, k, m)];
      }
  for (int m = 1; m <= 3; m++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        b[F4(i, j, 1, m)] = b[F4(i, j, nz - 1, m)];
        b[F4(i, j, nz, m)] = b[F4(i, j, 2, m)];
      }
}

/* periodic_scalar, boundary.f:167-206 */
void or_periodic_scalar(const orc *h, real *b) {
  int nx = h->nx, ny = h->ny, nz = h->nz;
  int sx1 = h->ox ? 2 : nx - 1, sxn = h->ox ? nx - 1 : 2;
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++) {
      b[F3(1, j, k)] = b[F3(sx1, j, k)];
      b[F3(nx, j, k)] = b[F3(sxn, j, k)];
    }
  for (int k = 1; k <= nz; k++)
    for (int i = 1; i <= nx; i++) {
      b[F3(i, 1, k)] = b[F3(i, ny - 1, k)];
      b[F3(i, ny, k)] = b[F3(i, 2, k)];
    }
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) {
      b[F3(i, j, 1)] = b[F3(i, j, nz - 1)];
      b[F3(i, j, nz)] = b[F3(i, j, 2)];
    }
}

/* face-centred B component -> node, the bilinear stencil shared by
   edge_to_center (grid_interp.f:22-55) and edge_to_face (:80-113) */
static void e2c_point(const orc *h, const real *bt, int i, int j, int k, real *o3) {
  int im = i - 1, jm = j - 1, km = k - 1;
  real b2, b1;
  b2 = bt[F4(i, jm, k, 1)] + h->yrat[j - 1] * (bt[F4(i, j, k, 1)] - bt[F4(i, jm, k, 1)]);
  b1 = bt[F4(i, jm, km, 1)] + h->yrat[j - 1] * (bt[F4(i, j, km, 1)] - bt[F4(i, jm, km, 1)]);
  o3[0] = b1 + h->zrat[k - 1] * (b2 - b1);
  b2 = bt[F4(im, j, k, 2)] + h->xrat[i - 1] * (bt[F4(i, j, k, 2)] - bt[F4(im, j, k, 2)]);
  b1 = bt[F4(im, j, km, 2)] + h->xrat[i - 1] * (bt[F4(i, j, km, 2)] - bt[F4(im, j, km, 2)]);
  o3[1] = b1 + h->zrat[k - 1] * (b2 - b1);
  b2 = bt[F4(i, jm, k, 3)] + h->yrat[j - 1] * (bt[F4(i, j, k, 3)] - bt[F4(i, jm, k, 3)]);
  b1 = bt[F4(im, jm, k, 3)] + h->yrat[j - 1] * (bt[F4(im, j, k, 3)] - bt[F4(im, jm, k, 3)]);
  o3[2] = b1 + h->xrat[i - 1] * (b2 - b1);
}

/* edge_to_center, grid_interp.f:10-62 */
void or_edge_to_center(const orc *h, real *bt, real *btc) {
  or_periodic(h, bt);
  OMP_FOR
  for (int k = 2; k <= h->nz - 1; k++)
    for (int j = 2; j <= h->ny - 1; j++)
      for (int i = 2; i <= h->nx - 1; i++) {
        real o[3];
        e2c_point(h, bt, i, j, k, o);
        for (int m = 1; m <= 3; m++) btc[F4(i, j, k, m)] = o[m - 1];
      }
  or_periodic(h, btc);
}

/* edge_to_face, grid_interp.f:67-130.  The reference loops i=2..nx and then
   reads btc(i+1) at i=nx, one past the array (grid_interp.f:119-127); those
   ghost-layer outputs are overwritten by the periodic() its only caller
   (face_to_center, via get_Ep1 gutsf.f:1267-1269) applies first, so the
   restatement computes the interior 2..n-1 and fills ghosts periodically. */
void or_edge_to_face(const orc *h, real *bt, real *btmf) {
  size_t n4 = NN * 3;
  real *btc = (real *)calloc(n4, sizeof(real));
  or_periodic(h, bt);
  OMP_FOR
  for (int k = 2; k <= h->nz; k++)
    for (int j = 2; j <= h->ny; j++)
      for (int i = 2; i <= h->nx; i++) {
        real o[3];
        e2c_point(h, bt, i, j, k, o);
        for (int m = 1; m <= 3; m++) btc[F4(i, j, k, m)] = o[m - 1];
      }
  or_periodic(h, btc);
  OMP_FOR
  for (int k = 2; k <= h->nz - 1; k++)
    for (int j = 2; j <= h->ny - 1; j++)
      for (int i = 2; i <= h->nx - 1; i++) {
        btmf[F4(i, j, k, 1)] = R(0.5) * (btc[F4(i, j, k, 1)] + btc[F4(i + 1, j, k, 1)]);
        btmf[F4(i, j, k, 2)] = R(0.5) * (btc[F4(i, j, k, 2)] + btc[F4(i, j + 1, k, 2)]);
        btmf[F4(i, j, k, 3)] = R(0.5) * (btc[F4(i, j, k, 3)] + btc[F4(i, j, k + 1, 3)]);
      }
  or_periodic(h, btmf);
  free(btc);
}

/* face_to_center, grid_interp.f:136-176 */
void or_face_to_center(const orc *h, real *v, real *vc) {
  or_periodic(h, v);
  OMP_FOR
  for (int k = 2; k <= h->nz; k++)
    for (int j = 2; j <= h->ny; j++)
      for (int i = 2; i <= h->nx; i++) {
        vc[F4(i, j, k, 1)] = h->xrat[i - 1] * (v[F4(i, j, k, 1)] - v[F4(i - 1, j, k, 1)]) + v[F4(i - 1, j, k, 1)];
        vc[F4(i, j, k, 2)] = h->yrat[j - 1] * (v[F4(i, j, k, 2)] - v[F4(i, j - 1, k, 2)]) + v[F4(i, j - 1, k, 2)];
        vc[F4(i, j, k, 3)] = h->zrat[k - 1] * (v[F4(i, j, k, 3)] - v[F4(i, j, k - 1, 3)]) + v[F4(i, j, k - 1, 3)];
      }
  or_periodic(h, vc);
}

/* crossf2, gutsf.f:40-104 (node cross product into module scratch ct, then
   the edge average) */
void or_crossf2(orc *h, real *aa, real *btc, real *cc) {
  real *ct = h->ct;
  or_periodic(h, aa);
  or_periodic(h, btc);
  OMP_FOR
  for (int k = 2; k <= h->nz - 1; k++)
    for (int j = 2; j <= h->ny - 1; j++)
      for (int i = 2; i <= h->nx - 1; i++) {
        real ax = aa[F4(i, j, k, 1)], bx = btc[F4(i, j, k, 1)];
        real ay = aa[F4(i, j, k, 2)], by = btc[F4(i, j, k, 2)];
        real az = aa[F4(i, j, k, 3)], bz = btc[F4(i, j, k, 3)];
        ct[F4(i, j, k, 1)] = ay * bz - az * by;
        ct[F4(i, j, k, 2)] = az * bx - ax * bz;
        ct[F4(i, j, k, 3)] = ax * by - ay * bx;
      }
  or_periodic(h, ct);
  OMP_FOR
  for (int k = 2; k <= h->nz - 1; k++)
    for (int j = 2; j <= h->ny - 1; j++)
      for (int i = 2; i <= h->nx - 1; i++) {
        cc[F4(i, j, k, 1)] = R(0.5) * (ct[F4(i, j, k, 1)] + ct[F4(i + 1, j, k, 1)]);
        cc[F4(i, j, k, 2)] = R(0.5) * (ct[F4(i, j, k, 2)] + ct[F4(i, j + 1, k, 2)]);
        cc[F4(i, j, k, 3)] = R(0.5) * (ct[F4(i, j, k, 3)] + ct[F4(i, j, k + 1, 3)]);
      }
  or_periodic(h, cc);
}

/* curlB, gutsf.f:178-244.  periodic(b1) mutates the input (gutsf.f:199);
   aj ghosts are not written (gutsf.f:241). */
void or_curlB(const orc *h, real *b1, const real *np, real *aj) {
  or_periodic(h, b1);
  OMP_FOR
  for (int k = 2; k <= h->nz - 1; k++)
    for (int j = 2; j <= h->ny - 1; j++)
      for (int i = 2; i <= h->nx - 1; i++) {
        real nt1 = R(0.5) * (np[F3(i, j, k)] + np[F3(i + 1, j, k)]);
        real nt2 = R(0.5) * (np[F3(i, j, k)] + np[F3(i, j + 1, k)]);
        real nt3 = R(0.5) * (np[F3(i, j, k)] + np[F3(i, j, k + 1)]);
        real dxc = h->dx_cell[i - 1], dyc = h->dy_cell[j - 1], dzc = h->dz_cell[k - 1];
        real c1 = (b1[F4(i, j, k, 3)] - b1[F4(i, j - 1, k, 3)]) / dyc +
                  (b1[F4(i, j, k - 1, 2)] - b1[F4(i, j, k, 2)]) / dzc;
        real c2 = (b1[F4(i, j, k, 1)] - b1[F4(i, j, k - 1, 1)]) / dzc +
                  (b1[F4(i - 1, j, k, 3)] - b1[F4(i, j, k, 3)]) / dxc;
        real c3 = (b1[F4(i, j, k, 2)] - b1[F4(i - 1, j, k, 2)]) / dxc +
                  (b1[F4(i, j - 1, k, 1)] - b1[F4(i, j, k, 1)]) / dyc;
        aj[F4(i, j, k, 1)] = c1 / (nt1 * h->alpha);
        aj[F4(i, j, k, 2)] = c2 / (nt2 * h->alpha);
        aj[F4(i, j, k, 3)] = c3 / (nt3 * h->alpha);
      }
}

/* curlE, gutsf.f:249-308 (interior only; callers never read its ghosts) */
void or_curlE(const orc *h, const real *E, real *cE) {
  OMP_FOR
  for (int i = 2; i <= h->nx - 1; i++)
    for (int j = 2; j <= h->ny - 1; j++)
      for (int k = 2; k <= h->nz - 1; k++) {
        real dxg = h->dx_grid[i - 1], dyg = h->dy_grid[j - 1], dzg = h->dz_grid[k - 1];
        cE[F4(i, j, k, 1)] = (E[F4(i, j + 1, k, 3)] - E[F4(i, j, k, 3)]) / dyg +
                             (E[F4(i, j, k, 2)] - E[F4(i, j, k + 1, 2)]) / dzg;
        cE[F4(i, j, k, 2)] = (E[F4(i, j, k, 3)] - E[F4(i + 1, j, k, 3)]) / dxg +
                             (E[F4(i, j, k + 1, 1)] - E[F4(i, j, k, 1)]) / dzg;
        cE[F4(i, j, k, 3)] = (E[F4(i, j, k, 1)] - E[F4(i, j + 1, k, 1)]) / dyg +
                             (E[F4(i + 1, j, k, 2)] - E[F4(i, j, k, 2)]) / dxg;
      }
}

/* E = c + nu*aj on the interior, then periodic (gutsf.f:1078-1089,
   :1278-1288) */
static void e_from_c(orc *h, real *E, const real *aj, const real *nu) {
  OMP_FOR
  for (int k = 2; k <= h->nz - 1; k++)
    for (int j = 2; j <= h->ny - 1; j++)
      for (int i = 2; i <= h->nx - 1; i++)
        for (int m = 1; m <= 3; m++)
          E[F4(i, j, k, m)] = h->c[F4(i, j, k, m)] + nu[F3(i, j, k)] * aj[F4(i, j, k, m)];
  or_periodic(h, E);
}

/* a = aj - up on the interior, into module scratch a (gutsf.f:1058-1064) */
static void a_from_aj_up(orc *h, const real *aj, const real *up) {
  OMP_FOR
  for (int k = 2; k <= h->nz - 1; k++)
    for (int j = 2; j <= h->ny - 1; j++)
      for (int i = 2; i <= h->nx - 1; i++)
        for (int m = 1; m <= 3; m++)
          h->a[F4(i, j, k, m)] = aj[F4(i, j, k, m)] - up[F4(i, j, k, m)];
}

/* get_E, gutsf.f:994-1110 */
void or_get_E(orc *h, real *E, real *bt, real *aj, real *up, const real *np,
              const real *nu) {
  (void)np;
  size_t n4 = NN * 3;
  real *aa = (real *)calloc(n4, sizeof(real));
  real *btc = (real *)calloc(n4, sizeof(real));
  a_from_aj_up(h, aj, up);
  or_face_to_center(h, h->a, aa);
  or_edge_to_center(h, bt, btc);
  or_crossf2(h, aa, btc, h->c);
  e_from_c(h, E, aj, nu);
  free(aa);
  free(btc);
}

/* predict_B, gutsf.f:1115-1169: b1p2 = b12 - 2 dtsub curlE */
void or_predict_B(orc *h, real *b1, real *b12, real *b1p2, real *bt, real *E,
                  real *aj, real *up, const real *np, const real *nu) {
  (void)b1;
  size_t n4 = NN * 3;
  real *cE = (real *)calloc(n4, sizeof(real));
  or_get_E(h, E, bt, aj, up, np, nu);
  or_curlE(h, E, cE);
  OMP_FOR
  for (int k = 2; k <= h->nz - 1; k++)
    for (int j = 2; j <= h->ny - 1; j++)
      for (int i = 2; i <= h->nx - 1; i++)
        for (int m = 1; m <= 3; m++)
          b1p2[F4(i, j, k, m)] = b12[F4(i, j, k, m)] - R(2.0) * h->dtsub * cE[F4(i, j, k, m)];
  or_periodic(h, b1p2);
  free(cE);
}

/* get_Ep1, gutsf.f:1174-1304: E at m+1/2 from btp1 = (b1p2+b1)/2 with the
   smoothed node field edge_to_face -> face_to_center */
void or_get_Ep1(orc *h, real *E, real *b1, real *b1p2, real *aj, real *up,
                const real *np, const real *nu) {
  size_t n4 = NN * 3;
  real *btp1 = (real *)calloc(n4, sizeof(real));
  real *btp1mf = (real *)calloc(n4, sizeof(real));
  real *btc = (real *)calloc(n4, sizeof(real));
  real *aa = (real *)calloc(n4, sizeof(real));
  OMP_FOR
  for (size_t n = 0; n < n4; n++) btp1[n] = R(0.5) * (b1p2[n] + b1[n]);
  or_curlB(h, btp1, np, aj);
  a_from_aj_up(h, aj, up);
  or_face_to_center(h, h->a, aa);
  or_edge_to_face(h, btp1, btp1mf);
  or_face_to_center(h, btp1mf, btc);
  or_crossf2(h, aa, btc, h->c);
  e_from_c(h, E, aj, nu);
  free(btp1); free(btp1mf); free(btc); free(aa);
}

/* correct_B, gutsf.f:1309-1359: b1p2 = b1 - dtsub curlE */
void or_correct_B(orc *h, real *b1, real *b1p2, real *E, real *aj, real *up,
                  const real *np, const real *nu) {
  size_t n4 = NN * 3;
  real *cE = (real *)calloc(n4, sizeof(real));
  or_get_Ep1(h, E, b1, b1p2, aj, up, np, nu);
  or_curlE(h, E, cE);
  OMP_FOR
  for (int k = 2; k <= h->nz - 1; k++)
    for (int j = 2; j <= h->ny - 1; j++)
      for (int i = 2; i <= h->nx - 1; i++)
        for (int m = 1; m <= 3; m++)
          b1p2[F4(i, j, k, m)] = b1[F4(i, j, k, m)] - h->dtsub * cE[F4(i, j, k, m)];
  or_periodic(h, b1p2);
  free(cE);
}

/* f_update_tlev, gutsf.f:10-35 (the B0 add is commented out there) */
void or_f_update_tlev(const orc *h, real *b1, real *b12, const real *b1p2, real *bt) {
  size_t n4 = NN * 3;
  OMP_FOR
  for (size_t n = 0; n < n4; n++) {
    bt[n] = b1p2[n];
    b12[n] = b1[n];
    b1[n] = b1p2[n];
  }
}

/* check_time_step, gutsf.f:1776-1822.  Node order is the reference's loop
   nest (i outer, j, k inner); every failing node halves dtsub once against
   the dtsub current at that node. */
void or_check_time_step(orc *h, const real *bt, const real *np) {
  for (int i = 1; i <= h->nx; i++)
    for (int j = 1; j <= h->ny; j++)
      for (int k = 1; k <= h->nz; k++) {
        real ak = R(2.) / h->dx;
        real b1 = bt[F4(i, j, k, 1)], b2 = bt[F4(i, j, k, 2)], b3 = bt[F4(i, j, k, 3)];
        real btot = (real)sqrt((double)(b1 * b1 + b2 * b2 + b3 * b3));
        real a1 = ak * ak * btot / (h->alpha * np[F3(i, j, k)]);
        real a2 = (ak * btot) * (ak * btot) / (h->alpha * np[F3(i, j, k)]);
        real womega = R(0.5) * (a1 + (real)sqrt((double)(a1 * a1 + R(4) * a2)));
        real phi = womega / ak;
        real deltat = h->dx / phi;
        if (deltat <= R(2.0) * h->dtsub) {
          h->dtsub = h->dtsub / R(2.0);
          h->ntf = h->ntf * R(2.0);
        }
      }
}

/* ---------------------------------------------------------------- */
/* particles (gutsp.f)                                               */
/* ---------------------------------------------------------------- */
#define P(arr, l, m) arr[(size_t)(Ni_max) * ((m)-1) + (size_t)((l)-1)]

/* smallest idx in 1..n with x <= q(idx), minus 1: the result of the
   reference's upward linear search (gutsp.f:1394-1401, 1407-1412), found by
   bisection (same answer for the monotone q arrays grd7 builds) */
static int search_cell(const real *q, int n, real x) {
  int lo = 1, hi = n; /* answer idx in [1, n] */
  if (x <= q[0]) return 0;
  while (lo < hi) {
    int mid = (lo + hi) / 2;
    if (x <= q[mid - 1]) hi = mid; else lo = mid + 1;
  }
  return lo - 1;
}

/* get_interp_weights, gutsp.f:1378-1435 */
void or_get_interp_weights(const orc *h, int Ni_tot, int Ni_max, const real *xp,
                           int32_t *ijkp, real *wght) {
  OMP_FOR
  for (int l = 1; l <= Ni_tot; l++) {
    real x = P(xp, l, 1), y = P(xp, l, 2), z = P(xp, l, 3);
    int i = search_cell(h->qx, h->nx, x);
    int j = (int)floor((double)(y / h->dy));
    int k = search_cell(h->qz, h->nz, z);
    P(ijkp, l, 1) = i; P(ijkp, l, 2) = j; P(ijkp, l, 3) = k;
    real vol = R(1.0) / ((QX(i + 1) - QX(i)) * (QY(j + 1) - QY(j)) * (QZ(k + 1) - QZ(k)));
    real x1 = (real)fabs((double)(x - QX(i))), x2 = (real)fabs((double)(x - QX(i + 1)));
    real y1 = (real)fabs((double)(y - QY(j))), y2 = (real)fabs((double)(y - QY(j + 1)));
    real z1 = (real)fabs((double)(z - QZ(k))), z2 = (real)fabs((double)(z - QZ(k + 1)));
    P(wght, l, 1) = x2 * y2 * z2 * vol;
    P(wght, l, 2) = x1 * y2 * z2 * vol;
    P(wght, l, 3) = x2 * y2 * z1 * vol;
    P(wght, l, 4) = x1 * y2 * z1 * vol;
    P(wght, l, 5) = x2 * y1 * z2 * vol;
    P(wght, l, 6) = x1 * y1 * z2 * vol;
    P(wght, l, 7) = x2 * y1 * z1 * vol;
    P(wght, l, 8) = x1 * y1 * z1 * vol;
  }
}

/* corner c (1..8) of the reference's weight order (gutsp.f:1422-1429) */
static inline void corner(int c, int i, int j, int k, int *ci, int *cj, int *ck) {
  static const int di[8] = {0, 1, 0, 1, 0, 1, 0, 1};
  static const int dj[8] = {0, 0, 0, 0, 1, 1, 1, 1};
  static const int dk[8] = {0, 0, 1, 1, 0, 0, 1, 1};
  *ci = i + di[c - 1]; *cj = j + dj[c - 1]; *ck = k + dk[c - 1];
}

/* fold ghost layer 1 into n-1 per axis (gutsp.f:1496-1498, 1803-1805).
   An open x folds ghost layer 1 onto layer 2 instead and then divides the
   two face layers by their share of the particle domain: layer 2 collects
   the whole first cell and half the second (1.5 cells), layer nx-1 half the
   last cell (0.5 cells). */
static void fold(const orc *h, real *g, int ncomp) {
  int nx = h->nx, ny = h->ny, nz = h->nz;
  int tx = h->ox ? 2 : nx - 1;
  for (int m = 1; m <= ncomp; m++) {
    real *s = g + NN * (size_t)(m - 1);
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++) {
        s[F3(tx, j, k)] += s[F3(1, j, k)];
        if (h->ox) {
          s[F3(1, j, k)] = 0;
          s[F3(2, j, k)] *= R(2.0) / R(3.0);
          s[F3(nx - 1, j, k)] *= R(2.0);
        }
      }
    for (int k = 1; k <= nz; k++)
      for (int i = 1; i <= nx; i++) s[F3(i, ny - 1, k)] += s[F3(i, 1, k)];
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) s[F3(i, j, nz - 1)] += s[F3(i, j, 1)];
  }
}

/* Deposit particles l0..l1 (1-based, inclusive) of update_np into g. */
static void np_deposit(const orc *h, int l0, int l1, int Ni_max, const int32_t *ijkp,
                       const real *wght, const real *beta_p, real *g) {
  for (int l = l0; l <= l1; l++) {
    int i = P(ijkp, l, 1), j = P(ijkp, l, 2), k = P(ijkp, l, 3);
    real bp = beta_p ? beta_p[l - 1] : R(1.0);
    real volb = R(1.0) / (h->dx_grid[i - 1] * h->dy_grid[j - 1] * h->dz_grid[k - 1] * h->beta * bp);
    for (int c = 1; c <= 8; c++) {
      int ci, cj, ck;
      corner(c, i, j, k, &ci, &cj, &ck);
      g[F3(ci, cj, ck)] += P(wght, l, c) * volb;
    }
  }
}

/* Particle deposits with h->nthreads > 1: thread t sums particles
   [1 + t N/T, (t+1) N/T] into its own zeroed copy of the ncomp-component grid,
   then out = the copies added in thread order. */
typedef void (*deposit_fn)(const orc *h, int l0, int l1, const void *args, real *g);
static void deposit_threads(const orc *h, int N, int ncomp, deposit_fn fn, const void *args,
                            real *out) {
  const int T = h->nthreads;
  const size_t len = NN * (size_t)ncomp;
  real *priv = (real *)calloc((size_t)T * len, sizeof(real));
#pragma omp parallel num_threads(T)
  {
#ifdef _OPENMP
    const int t = omp_get_thread_num();
#else
    const int t = 0;
#endif
    const int l0 = 1 + (int)((int64_t)t * N / T), l1 = (int)((int64_t)(t + 1) * N / T);
    fn(h, l0, l1, args, priv + (size_t)t * len);
  }
#pragma omp parallel for schedule(static) num_threads(T)
  for (size_t q = 0; q < len; q++) {
    real a = priv[q];
    for (int t = 1; t < T; t++) a += priv[(size_t)t * len + q];
    out[q] = a;
  }
  free(priv);
}

typedef struct {
  int Ni_max;
  const int32_t *ijkp;
  const real *wght, *beta_p, *vp, *np;
} dep_args;
static void np_deposit_fn(const orc *h, int l0, int l1, const void *a, real *g) {
  const dep_args *d = (const dep_args *)a;
  np_deposit(h, l0, l1, d->Ni_max, d->ijkp, d->wght, d->beta_p, g);
}

/* update_np, gutsp.f:1441-1512 (one rank: the allreduce is the identity) */
void or_update_np(orc *h, int Ni_tot, int Ni_max, const int32_t *ijkp,
                  const real *wght, const real *beta_p, real *np) {
  memset(np, 0, sizeof(real) * NN);
  if (h->nthreads > 1) {
    dep_args a = {Ni_max, ijkp, wght, beta_p, NULL, NULL};
    deposit_threads(h, Ni_tot, 1, np_deposit_fn, &a, np);
  } else {
    np_deposit(h, 1, Ni_tot, Ni_max, ijkp, wght, beta_p, np);
  }
  fold(h, np, 1);
  or_periodic_scalar(h, np);
}

/* Deposit particles l0..l1 of update_up into ct (gutsp.f:1704-1756: a
   division by np per particle and corner) */
static void up_deposit(const orc *h, int l0, int l1, int Ni_max, const int32_t *ijkp,
                       const real *wght, const real *beta_p, const real *vp, const real *np,
                       real *ct) {
  for (int l = l0; l <= l1; l++) {
    int i = P(ijkp, l, 1), j = P(ijkp, l, 2), k = P(ijkp, l, 3);
    real bp = beta_p ? beta_p[l - 1] : R(1.0);
    real volb = (QX(i + 1) - QX(i)) * (QY(j + 1) - QY(j)) * (QZ(k + 1) - QZ(k)) * h->beta * bp;
    for (int c = 1; c <= 8; c++) {
      int ci, cj, ck;
      corner(c, i, j, k, &ci, &cj, &ck);
      real nvolb = R(1.0) / (np[F3(ci, cj, ck)] * volb);
      for (int m = 1; m <= 3; m++)
        ct[F4(ci, cj, ck, m)] += P(vp, l, m) * P(wght, l, c) * nvolb;
    }
  }
}
static void up_deposit_fn(const orc *h, int l0, int l1, const void *a, real *g) {
  const dep_args *d = (const dep_args *)a;
  up_deposit(h, l0, l1, d->Ni_max, d->ijkp, d->wght, d->beta_p, d->vp, d->np, g);
}

/* update_up, gutsp.f:1649-1842 */
void or_update_up(orc *h, int Ni_tot, int Ni_max, const int32_t *ijkp,
                  const real *wght, const real *beta_p, const real *vp, real *np,
                  real *up) {
  real *ct = h->ct;
  memset(up, 0, sizeof(real) * NN * 3);
  memset(ct, 0, sizeof(real) * NN * 3);
  if (h->nthreads > 1) {
    dep_args a = {Ni_max, ijkp, wght, beta_p, vp, np};
    deposit_threads(h, Ni_tot, 3, up_deposit_fn, &a, ct);
  } else {
    up_deposit(h, 1, Ni_tot, Ni_max, ijkp, wght, beta_p, vp, np, ct);
  }
  fold(h, ct, 3);
  or_periodic(h, ct);
  OMP_FOR
  for (int k = 1; k <= h->nz - 1; k++)
    for (int j = 1; j <= h->ny - 1; j++)
      for (int i = 1; i <= h->nx - 1; i++) {
        up[F4(i, j, k, 1)] = R(0.5) * (ct[F4(i, j, k, 1)] + ct[F4(i + 1, j, k, 1)]);
        up[F4(i, j, k, 2)] = R(0.5) * (ct[F4(i, j, k, 2)] + ct[F4(i, j + 1, k, 2)]);
        up[F4(i, j, k, 3)] = R(0.5) * (ct[F4(i, j, k, 3)] + ct[F4(i, j, k + 1, 3)]);
      }
  or_periodic(h, up);
}

/* extrapol_up, gutsp.f:127-150: deposit v = 1.5 vp - 0.5 vp1 */
void or_extrapol_up(orc *h, int Ni_tot, int Ni_max, const int32_t *ijkp,
                    const real *wght, const real *beta_p, const real *vp,
                    const real *vp1, real *np, real *up) {
  real *v = (real *)malloc(sizeof(real) * (size_t)Ni_max * 3);
  for (int m = 1; m <= 3; m++) {
    OMP_FOR
    for (int l = 1; l <= Ni_tot; l++) P(v, l, m) = R(1.5) * P(vp, l, m) - R(0.5) * P(vp1, l, m);
  }
  or_update_up(h, Ni_tot, Ni_max, ijkp, wght, beta_p, v, np, up);
  free(v);
}

/* trilinear gather of a vector grid at particle l (gutsp.f:217-251) */
static inline real gather(const orc *h, const real *g, int m, int i, int j, int k,
                          const real *w8) {
  int ip = i + 1, jp = j + 1, kp = k + 1;
  return g[F4(i, j, k, m)] * w8[0] + g[F4(ip, j, k, m)] * w8[1] + g[F4(i, j, kp, m)] * w8[2] +
         g[F4(ip, j, kp, m)] * w8[3] + g[F4(i, jp, k, m)] * w8[4] + g[F4(ip, jp, k, m)] * w8[5] +
         g[F4(i, jp, kp, m)] * w8[6] + g[F4(ip, jp, kp, m)] * w8[7];
}

/* get_Ep, gutsp.f:155-276 */
void or_get_Ep(orc *h, int Ni_tot, int Ni_max, const int32_t *ijkp, const real *wght,
               const real *mrat, real *aj, real *up, const real *btc, real *Ep) {
  size_t n4 = NN * 3;
  real *ajc = (real *)calloc(n4, sizeof(real));
  real *upc = (real *)calloc(n4, sizeof(real));
  or_face_to_center(h, aj, ajc);
  or_face_to_center(h, up, upc);
  OMP_FOR
  for (int l = 1; l <= Ni_tot; l++) {
    int i = P(ijkp, l, 1), j = P(ijkp, l, 2), k = P(ijkp, l, 3);
    real w8[8];
    for (int c = 0; c < 8; c++) w8[c] = P(wght, l, c + 1);
    real aa[3], bb[3];
    for (int m = 1; m <= 3; m++) {
      real aj3 = gather(h, ajc, m, i, j, k, w8);
      real up3 = gather(h, upc, m, i, j, k, w8);
      aa[m - 1] = aj3 - up3;
      bb[m - 1] = gather(h, btc, m, i, j, k, w8);
    }
    real mr = mrat ? mrat[l - 1] : R(1.0);
    P(Ep, l, 1) = (aa[1] * bb[2] - aa[2] * bb[1]) * mr;
    P(Ep, l, 2) = (aa[2] * bb[0] - aa[0] * bb[2]) * mr;
    P(Ep, l, 3) = (aa[0] * bb[1] - aa[1] * bb[0]) * mr;
  }
  free(ajc);
  free(upc);
}

/* get_vplus_vminus, gutsp.f:281-402 (the Crank-Nicolson rotation) */
void or_get_vplus_vminus(const orc *h, int Ni_tot, int Ni_max, const int32_t *ijkp,
                         const real *wght, const real *mrat, const real *Ep,
                         const real *btc, const real *vp, real *vplus, real *vminus) {
  real dt = h->dt;
  for (int m = 1; m <= 3; m++) {
    OMP_FOR
    for (int l = 1; l <= Ni_tot; l++) P(vminus, l, m) = P(vp, l, m) + R(0.5) * dt * P(Ep, l, m);
  }
  OMP_FOR
  for (int l = 1; l <= Ni_tot; l++) {
    int i = P(ijkp, l, 1), j = P(ijkp, l, 2), k = P(ijkp, l, 3);
    real w8[8];
    for (int c = 0; c < 8; c++) w8[c] = P(wght, l, c + 1);
    real b3[3];
    for (int m = 1; m <= 3; m++) b3[m - 1] = gather(h, btc, m, i, j, k, w8);
    real mr = mrat ? mrat[l - 1] : R(1.0);
    real vm1 = P(vminus, l, 1), vm2 = P(vminus, l, 2), vm3 = P(vminus, l, 3);
    real vxb[3];
    vxb[0] = vm2 * b3[2] * mr - vm3 * b3[1] * mr;
    vxb[1] = vm3 * b3[0] * mr - vm1 * b3[2] * mr;
    vxb[2] = vm1 * b3[1] * mr - vm2 * b3[0] * mr;
    real vdb = vm1 * b3[0] * mr + vm2 * b3[1] * mr + vm3 * b3[2] * mr;
    real Bx = b3[0] * mr, By = b3[1] * mr, Bz = b3[2] * mr;
    real B2 = Bx * Bx + By * By + Bz * Bz;
    real dt2 = dt * dt;
    real a_d = R(1.0) / (R(1.0) + (R(0.25) * B2 * dt2));
    real a1 = (R(1) - (R(0.25) * B2 * dt2)) * a_d;
    real a2 = dt * a_d;
    real a3 = R(0.5) * dt2 * a_d;
    for (int m = 1; m <= 3; m++)
      P(vplus, l, m) = a1 * P(vminus, l, m) + a2 * vxb[m - 1] + a3 * vdb * b3[m - 1] * mr;
  }
}

/* improve_up, gutsp.f:407-428 */
void or_improve_up(orc *h, int Ni_tot, int Ni_max, const int32_t *ijkp, const real *wght,
                   const real *beta_p, real *vp1, const real *vplus, const real *vminus,
                   real *np, real *up) {
  for (int m = 1; m <= 3; m++) {
    OMP_FOR
    for (int l = 1; l <= Ni_tot; l++)
      P(vp1, l, m) = R(0.5) * (P(vplus, l, m) + P(vminus, l, m));
  }
  or_update_up(h, Ni_tot, Ni_max, ijkp, wght, beta_p, vp1, np, up);
}

/* get_vp_final, gutsp.f:433-449 */
void or_get_vp_final(const orc *h, int Ni_tot, int Ni_max, const real *Ep, real *vp,
                     real *vp1, const real *vplus) {
  for (int m = 1; m <= 3; m++) {
    OMP_FOR
    for (int l = 1; l <= Ni_tot; l++) {
      P(vp1, l, m) = P(vp, l, m);
      P(vp, l, m) = P(vplus, l, m) + R(0.5) * h->dt * P(Ep, l, m);
    }
  }
}

/* move_ion_half, gutsp.f:454-570: x += dt/2 v, one periodic wrap per axis
   into (q(1), q(n-1)]; an open x is not wrapped (or_remove_escaped then
   drops the ions outside) */
void or_move_ion_half(const orc *h, int Ni_tot, int Ni_max, real *xp, const real *vp) {
  real dth = h->dt / R(2.);
  const real *q[3] = {h->qx, h->qy, h->qz};
  int n[3] = {h->nx, h->ny, h->nz};
  OMP_FOR
  for (int l = 1; l <= Ni_tot; l++)
    for (int m = 1; m <= 3; m++) {
      real x = P(xp, l, m) + dth * P(vp, l, m);
      real q1 = q[m - 1][0], qn = q[m - 1][n[m - 1] - 2];
      if (!(m == 1 && h->ox)) {
        if (x > qn) x = q1 + (x - qn);
        else if (x <= q1) x = qn - (q1 - x);
      }
      P(xp, l, m) = x;
    }
}

/* an open x (build-defined): drop the ions outside (q(1), q(nx-1)] in x,
   keeping the order of the rest (remove_ion, gutsp.f:12-46, one ion at a
   time there); returns the count removed */
int or_remove_escaped(const orc *h, or_state *s) {
  const int N = s->Ni_tot, M = s->Ni_max;
  const real q1 = h->qx[0], qn = h->qx[h->nx - 2];
  int w = 0;
  for (int l = 0; l < N; l++) {
    real x = s->xp[l];
    if (x <= q1 || x > qn) {
      s->input_k -= or_ke(h, s, l + 1, l + 2);
      continue;
    }
    for (int m = 0; m < 3; m++) {
      s->xp[(size_t)m * M + w] = s->xp[(size_t)m * M + l];
      s->vp[(size_t)m * M + w] = s->vp[(size_t)m * M + l];
      s->vp1[(size_t)m * M + w] = s->vp1[(size_t)m * M + l];
    }
    if (s->beta_p) s->beta_p[w] = s->beta_p[l];
    if (s->mrat) s->mrat[w] = s->mrat[l];
    if (s->id) s->id[w] = s->id[l];
    w++;
  }
  s->Ni_tot = w;
  return N - w;
}

static void inflow_append(const orc *h, or_state *s, const double *t, const double *v,
                          int64_t cap, int64_t c, uint32_t id);

/* one step in the order of maind.f:385-479 (ionization and the energy
   diagnostic of maind.f:379-393 left out) */
void or_step(orc *h, or_state *s) {
  int N = s->Ni_tot, M = s->Ni_max;
  s->escaped = 0;
  s->injected = 0;
  /* an open x with inflow: this step's candidates (swap_set_inflow) */
  int64_t nin = 0, cap = 0;
  double *th = NULL, *te = NULL, *vin = NULL;
  uint8_t *flag = NULL;
  uint32_t *ids = NULL;
  if (h->ox && h->in_ppc > 0) {
    const int nc[3] = {h->nx - 2, h->ny - 2, h->nz - 2};
    const double d[3] = {h->dx, h->dy, h->delz};
    cap = (int64_t)(h->in_ppc * (h->nx - 2) * (h->ny - 2) * (h->nz - 2)) + 1024;
    th = (double *)malloc(sizeof(double) * 3 * cap);
    te = (double *)malloc(sizeof(double) * 3 * cap);
    vin = (double *)malloc(sizeof(double) * 3 * cap);
    flag = (uint8_t *)malloc((size_t)cap);
    ids = (uint32_t *)malloc(sizeof(uint32_t) * cap);
    nin = or_inflow(h->in_seed, (uint32_t)h->step, nc, h->dt, d, h->in_ppc, h->in_vsw, h->in_vth,
                    th, te, vin, flag, cap);
    if (nin > cap) nin = cap;
    for (int64_t c = 0; c < nin; c++) ids[c] = (flag[c] & 2) ? h->next_id++ : 0xFFFFFFFFu;
  }
  or_get_interp_weights(h, N, M, s->xp, s->ijkp, s->wght);
  or_update_np(h, N, M, s->ijkp, s->wght, s->beta_p, s->np);
  or_update_up(h, N, M, s->ijkp, s->wght, s->beta_p, s->vp, s->np, s->up);
  or_curlB(h, s->bt, s->np, s->aj);
  or_edge_to_center(h, s->bt, s->btc);
  or_extrapol_up(h, N, M, s->ijkp, s->wght, s->beta_p, s->vp, s->vp1, s->np, s->up);
  or_get_Ep(h, N, M, s->ijkp, s->wght, s->mrat, s->aj, s->up, s->btc, s->Ep);
  or_get_vplus_vminus(h, N, M, s->ijkp, s->wght, s->mrat, s->Ep, s->btc, s->vp, s->vplus, s->vminus);
  or_improve_up(h, N, M, s->ijkp, s->wght, s->beta_p, s->vp1, s->vplus, s->vminus, s->np, s->up);
  or_get_Ep(h, N, M, s->ijkp, s->wght, s->mrat, s->aj, s->up, s->btc, s->Ep);
  or_get_vplus_vminus(h, N, M, s->ijkp, s->wght, s->mrat, s->Ep, s->btc, s->vp, s->vplus, s->vminus);
  or_get_vp_final(h, N, M, s->Ep, s->vp, s->vp1, s->vplus);
  or_move_ion_half(h, N, M, s->xp, s->vp);
  if (h->ox) {
    s->escaped += or_remove_escaped(h, s);
    /* candidates inside at x^{n+1/2} join before the deposit */
    for (int64_t c = 0; c < nin; c++)
      if (flag[c] & 1) inflow_append(h, s, th, vin, cap, c, ids[c]);
    N = s->Ni_tot;
  }
  or_get_interp_weights(h, N, M, s->xp, s->ijkp, s->wght);
  or_update_np(h, N, M, s->ijkp, s->wght, s->beta_p, s->np);
  or_update_up(h, N, M, s->ijkp, s->wght, s->beta_p, s->vp, s->np, s->up);
  h->dtsub = h->dt / (real)h->ntsub; /* maind.f:420-421 */
  h->ntf = (real)h->ntsub;
  or_check_time_step(h, s->bt, s->np);
  for (int n = 1; n <= (int)h->ntf; n++) {
    or_curlB(h, s->bt, s->np, s->aj);
    or_predict_B(h, s->b1, s->b12, s->b1p2, s->bt, s->E, s->aj, s->up, s->np, s->nu);
    or_correct_B(h, s->b1, s->b1p2, s->E, s->aj, s->up, s->np, s->nu);
    or_f_update_tlev(h, s->b1, s->b12, s->b1p2, s->bt);
  }
  or_move_ion_half(h, N, M, s->xp, s->vp);
  if (h->ox) {
    /* candidates that entered in the first half and left in the second go
       with the escapes but are neither injected nor escaped ions */
    int passing = 0;
    for (int64_t c = 0; c < nin; c++) passing += flag[c] == 1;
    s->escaped += or_remove_escaped(h, s) - passing;
    for (int64_t c = 0; c < nin; c++)
      if ((flag[c] & 2) && !(flag[c] & 1)) inflow_append(h, s, te, vin, cap, c, ids[c]);
    for (int64_t c = 0; c < nin; c++) s->injected += (flag[c] & 2) ? 1 : 0;
  }
  free(th); free(te); free(vin); free(flag); free(ids);
  h->step++;
}

/* ---------------------------------------------------------------- */
/* build-only pieces                                                 */
/* ---------------------------------------------------------------- */

/* Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
   easy as 1, 2, 3", SC'11): 10 rounds of two 32x32->64 multiplies with the
   published multipliers and Weyl key increments. */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* split-precision half move; the device kernel's arithmetic, written with
   explicit fmaf so both sides round identically */
void or_split_move(int64_t n, uint32_t *cell, float *fx, float *fy, float *fz,
                   const float *vx, const float *vy, const float *vz, const float s[3],
                   const int nc[3]) {
  for (int64_t p = 0; p < n; p++) {
    uint32_t c = cell[p];
    int ijk[3] = {(int)(c & 2047u), (int)((c >> 11) & 1023u), (int)(c >> 21)};
    float *f[3] = {&fx[p], &fy[p], &fz[p]};
    float v[3] = {vx[p], vy[p], vz[p]};
    for (int a = 0; a < 3; a++) {
      float t = fmaf(v[a], s[a], *f[a]);
      float sh = ceilf(t) - 1.0f; /* right-closed cells: f in (0,1] */
      *f[a] = t - sh;
      int ci = ijk[a] + (int)sh;
      ci %= nc[a];
      if (ci < 0) ci += nc[a];
      ijk[a] = ci;
    }
    cell[p] = (uint32_t)ijk[0] | ((uint32_t)ijk[1] << 11) | ((uint32_t)ijk[2] << 21);
  }
}

void or_stable_sort(int64_t n, const uint32_t *key, uint32_t nkeys, uint32_t *perm) {
  uint32_t *cnt = (uint32_t *)calloc((size_t)nkeys + 1, sizeof(uint32_t));
  for (int64_t p = 0; p < n; p++) cnt[key[p] + 1]++;
  for (uint32_t k = 0; k < nkeys; k++) cnt[k + 1] += cnt[k];
  for (int64_t p = 0; p < n; p++) perm[cnt[key[p]]++] = (uint32_t)p;
  free(cnt);
}

/* the drifting-Maxwellian sampler of load_Maxwellian (part_init.f:176-206)
   as the device loader draws it: Philox counter (id lo, id hi, stream, draw
   0..2) keyed by the 64-bit seed; position t = (1-u) n_cells in cell units
   (exact in double), right-closed split t = i + f with f in (0,1]; velocity
   vsw + vth sqrt(-ln u1) cos(PI u2) with PI = 3.14159 (inputs.f:52). */
static double or_u24(uint32_t r) { return (double)(r >> 8) * (1.0 / 16777216.0); }
void or_load_maxwellian(int64_t n, uint64_t seed, uint32_t stream, const int nc[3], double vsw,
                        double vth, uint32_t *cell, double *frac, double *v) {
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int64_t p = 0; p < n; p++) {
    uint32_t r[3][4];
    for (int d = 0; d < 3; d++) {
      uint32_t ctr[4] = {(uint32_t)p, (uint32_t)((uint64_t)p >> 32), stream, (uint32_t)d};
      or_philox4x32_10(ctr, key, r[d]);
    }
    int ijk[3];
    for (int a = 0; a < 3; a++) {
      double t = (1.0 - or_u24(r[0][a])) * (double)nc[a];
      double i = ceil(t) - 1.0;
      ijk[a] = (int)i;
      frac[a * n + p] = t - i;
    }
    cell[p] = (uint32_t)ijk[0] | ((uint32_t)ijk[1] << 11) | ((uint32_t)ijk[2] << 21);
    const uint32_t *ur[6] = {&r[0][3], &r[1][0], &r[1][1], &r[1][2], &r[1][3], &r[2][0]};
    for (int a = 0; a < 3; a++) {
      double u1 = 1.0 - or_u24(*ur[2 * a]);
      double u2 = or_u24(*ur[2 * a + 1]);
      v[a * n + p] = (a == 0 ? vsw : 0.0) + vth * sqrt(-log(u1)) * cos(3.14159 * u2);
    }
  }
}

/* Ionize_sw_mp (chem_rates.f:508-646) with the build's RNG in place of
   pad_ranf.  The count (:553-562): dNi = ndot dt beta beta_pu dx dy delz nz
   with nz the reference's z extent including its two ghost layers (the
   formula has no nx, ny: the reference's quasi-1-D volume); below one ion
   dNi becomes 1 when dNi > u (Philox counter (step, 0, 0x10, 0), word 0),
   then nint.  The new ions (:567-610): ion p = 0..dNi-1 draws Philox counter
   (p lo, p hi, 0x20000 + step, 7); word 0 gives the ring phase
   theta = u 2 PI (PI = 3.14159, inputs.f:52) and v = (vsw + vsw cos theta,
   vsw sin theta, 0) (:570-573); words 1..3 give x_a = q(1) + (1 - u)(q(n-1) -
   q(1)) (:575-577), i.e. t = (1 - u) nc_a cells above q(1), split
   right-closed as get_interp_weights' search (:579-593) does: cell
   ceil(t) - 1, fraction t - cell in (0, 1].  mrat = 1 and beta_p = beta_pu
   (:606-608) are the caller's; the re-deposit of :641-643 is the caller's
   next update_np/update_up.  Returns dNi; writes at most cap ions
   (cell packed i | j<<11 | k<<21, frac and v as [3][cap]). */
int64_t or_ionize_sw(const int nc[3], double dx, double dy, double dz, double dt, double beta,
                     double ndot, double vsw, double beta_pu, uint64_t seed, uint32_t step,
                     uint32_t *cell, double *frac, double *v, int64_t cap) {
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  double dNi = ndot * dt * beta * beta_pu * dx * dy * dz * (double)(nc[2] + 2);
  if (dNi < 1.0) {
    uint32_t ctr[4] = {step, 0u, 0x10u, 0u}, r[4];
    or_philox4x32_10(ctr, key, r);
    if (dNi > or_u24(r[0])) dNi = 1.0;
  }
  int64_t m = (int64_t)llround(dNi);
  if (m < 0) m = 0;
  for (int64_t p = 0; p < m && p < cap; p++) {
    uint32_t ctr[4] = {(uint32_t)p, (uint32_t)((uint64_t)p >> 32), 0x20000u + step, 7u}, r[4];
    or_philox4x32_10(ctr, key, r);
    double th = or_u24(r[0]) * 2.0 * 3.14159;
    int ijk[3];
    for (int a = 0; a < 3; a++) {
      double t = (1.0 - or_u24(r[1 + a])) * (double)nc[a];
      double i = ceil(t) - 1.0;
      ijk[a] = (int)i;
      frac[a * cap + p] = t - i;
    }
    cell[p] = (uint32_t)ijk[0] | ((uint32_t)ijk[1] << 11) | ((uint32_t)ijk[2] << 21);
    v[p] = vsw + vsw * cos(th);
    v[cap + p] = vsw * sin(th);
    v[2 * cap + p] = 0.0;
  }
  return m;
}

/* pickup-ion cloud: photoionization counts and new ions per interior node,
   and charge exchange, following Ionize_pluto_mp (chem_rates.f:290-503) and
   res_chex (chem_rates.f:805-843) with the neutral profile
   n_n = N0 (r0/r)^2 exp(-r/H), r >= rmin.  cl = {cx, cy, cz, N0, r0, H,
   tau_photo, rmin, sigma_chex, mrat_pu, beta_min}; grid positions q(i) = i d.
   A node with more than max_per_cell new ions gets m = max(max_per_cell,
   ceil(beta_min n)) ions (at most n) of weight beta_p = m/n. */
static double or_cloud_nn(const double *cl, double x, double y, double z) {
  double r = sqrt((x - cl[0]) * (x - cl[0]) + (y - cl[1]) * (y - cl[1]) +
                  (z - cl[2]) * (z - cl[2]));
  r = r > cl[7] ? r : cl[7];
  double a = cl[4] / r;
  return cl[3] * a * a * exp(-r / cl[5]);
}

void or_cloud_counts(const int nc[3], double dx, double dy, double dz, double beta, double dt,
                     const double *cl, int max_per_cell, uint64_t seed, uint32_t step,
                     uint32_t *cnt, float *bp) {
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  double rate = dx * dy * dz * beta * dt / cl[6]; /* chem_rates.f:352 */
  int64_t q = 0;
  for (int k = 1; k <= nc[2]; k++)
    for (int j = 1; j <= nc[1]; j++)
      for (int i = 1; i <= nc[0]; i++, q++) {
        double npofr = rate * or_cloud_nn(cl, (i + 1) * dx, (j + 1) * dy, (k + 1) * dz);
        uint32_t n;
        if (npofr >= 1.0) {
          n = npofr < 4.0e9 ? (uint32_t)llround(npofr) : 4000000000u; /* :354 */
        } else {
          uint32_t ctr[4] = {(uint32_t)q, (uint32_t)((uint64_t)q >> 32), 0x38000u + step,
                             0xFFFFFFFFu}, r[4];
          or_philox4x32_10(ctr, key, r);
          n = npofr > or_u24(r[0]) ? 1u : 0u; /* :391 */
        }
        float w = 1.0f;
        if (n > (uint32_t)max_per_cell) {
          double mm = (double)max_per_cell, lo = ceil(cl[10] * (double)n);
          if (lo > mm) mm = lo < (double)n ? lo : (double)n;
          w = (float)(mm / (double)n);
          n = (uint32_t)mm;
        }
        cnt[q] = n;
        bp[q] = w;
      }
}

void or_cloud_ions(const int nc[3], const uint32_t *cnt, uint64_t seed, uint32_t step,
                   int open_x, uint32_t *cell, double *frac, int64_t m) {
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  int64_t q = 0, p = 0;
  for (int k = 1; k <= nc[2]; k++)
    for (int j = 1; j <= nc[1]; j++)
      for (int i = 1; i <= nc[0]; i++, q++)
        for (uint32_t l = 0; l < cnt[q]; l++, p++) {
          uint32_t ctr[4] = {(uint32_t)q, (uint32_t)((uint64_t)q >> 32), 0x30000u + step, l},
                   r[4];
          or_philox4x32_10(ctr, key, r);
          const int s[3] = {i, j, k};
          int ijk[3];
          for (int a = 0; a < 3; a++) {
            double t = (double)s[a] + or_u24(r[a]) - 0.5; /* chem_rates.f:363-365 */
            /* past the last face: the periodic image, or the mirror image
               on an open x */
            if (t > (double)nc[a]) t = a == 0 && open_x ? 2.0 * nc[a] - t : t - (double)nc[a];
            double ci = ceil(t) - 1.0;
            ijk[a] = (int)ci;
            frac[a * m + p] = t - ci;
          }
          cell[p] = (uint32_t)ijk[0] | ((uint32_t)ijk[1] << 11) | ((uint32_t)ijk[2] << 21);
        }
}

void or_chex(int64_t n, const uint32_t *cell, const float *fx, const float *fy, const float *fz,
             const float *vx, const float *vy, const float *vz, const uint32_t *id, double dx,
             double dy, double dz, double dt, const double *cl, uint64_t seed, uint32_t step,
             uint8_t *hit) {
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int64_t p = 0; p < n; p++) {
    uint32_t c = cell[p];
    double x = ((double)(c & 2047u) + 1.0 + (double)fx[p]) * dx;
    double y = ((double)((c >> 11) & 1023u) + 1.0 + (double)fy[p]) * dy;
    double z = ((double)(c >> 21) + 1.0 + (double)fz[p]) * dz;
    double a = vx[p], b = vy[p], e = vz[p];
    double prob = dt * or_cloud_nn(cl, x, y, z) * cl[8] * sqrt(a * a + b * b + e * e);
    const uint64_t who = id ? (uint64_t)id[p] : (uint64_t)p; /* the ion's identity */
    uint32_t ctr[4] = {(uint32_t)who, (uint32_t)(who >> 32), 0x40000u + step, 0u}, r[4];
    or_philox4x32_10(ctr, key, r);
    hit[p] = or_u24(r[0]) < prob ? 1 : 0;
  }
}

/* the inflow candidates of one step (build-defined inflow through the open
   x face; cf. check_min_den_boundary gutsp.f:575-816, which refills a
   boundary layer instead): slab depth L = max(1, ceil((|vsw| + 4.08 vth)
   dt/dx)) cells, spc = round(ppc L) candidates per y-z column of the face
   and nslot = spc ncy ncz; slot s belongs to column c = s / spc (y index
   c % ncy, z index c / ncy), draws Philox counter (s, step, 0x50000, d)
   d = 0..2 keyed by the seed, start position x = (1-u) L - L,
   y = jy + (1-u), z = jz + (1-u) (cell units, inside its column),
   velocity as the loader, then two half moves t += v dt/(2 d) (explicit fma;
   y and z wrapped into (0, nc]).  flag bit 0: inside after the first half
   move (deposits at x^{n+1/2}); bit 1: inside after the second and not past
   the far face after the first (appended).  Writes th, te, v ([3][cap]) and
   flag of the flagged slots in slot order; returns their count (at most cap
   written), or -1 for bad parameters. */
static double or_wrap_open(double t, double n) {
  if (t > n || t <= 0.0) {
    t -= (ceil(t / n) - 1.0) * n;
    if (t <= 0.0) t += n;
    else if (t > n) t -= n;
  }
  return t;
}
int64_t or_inflow(uint64_t seed, uint32_t step, const int nc[3], double dt, const double d[3],
                  double ppc, double vsw, double vth, double *th, double *te, double *v,
                  uint8_t *flag, int64_t cap) {
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  double sh[3] = {0.5 * dt / d[0], 0.5 * dt / d[1], 0.5 * dt / d[2]};
  double reach = (fabs(vsw) + 4.08 * vth) * dt / d[0];
  if (!(reach < 1048576.0)) return -1;
  int L = (int)ceil(reach);
  if (L < 1) L = 1;
  double spc = round(ppc * (double)L);
  double ns = spc * (double)nc[1] * (double)nc[2];
  if (!(ns < 4.0e9)) return -1;
  uint32_t nslot = (uint32_t)ns;
  int64_t m = 0;
  for (uint32_t q = 0; q < nslot; q++) {
    uint32_t r[3][4];
    for (int w = 0; w < 3; w++) {
      uint32_t ctr[4] = {q, step, 0x50000u, (uint32_t)w};
      or_philox4x32_10(ctr, key, r[w]);
    }
    const uint32_t *ur[6] = {&r[0][3], &r[1][0], &r[1][1], &r[1][2], &r[1][3], &r[2][0]};
    double vv[3], t0[3], a1[3], a2[3];
    for (int a = 0; a < 3; a++) {
      double u1 = 1.0 - or_u24(*ur[2 * a]);
      double u2 = or_u24(*ur[2 * a + 1]);
      vv[a] = (a == 0 ? vsw : 0.0) + vth * sqrt(-log(u1)) * cos(3.14159 * u2);
    }
    t0[0] = (1.0 - or_u24(r[0][0])) * (double)L - (double)L;
    uint32_t col = q / (uint32_t)spc;
    t0[1] = (double)(col % (uint32_t)nc[1]) + (1.0 - or_u24(r[0][1]));
    t0[2] = (double)(col / (uint32_t)nc[1]) + (1.0 - or_u24(r[0][2]));
    for (int a = 0; a < 3; a++) {
      a1[a] = fma(vv[a], sh[a], t0[a]);
      a2[a] = fma(vv[a], sh[a], a1[a]);
    }
    for (int a = 1; a < 3; a++) {
      a1[a] = or_wrap_open(a1[a], (double)nc[a]);
      a2[a] = or_wrap_open(a2[a], (double)nc[a]);
    }
    double ncx = (double)nc[0];
    int dep = a1[0] > 0.0 && a1[0] <= ncx;
    int keep = a2[0] > 0.0 && a2[0] <= ncx && a1[0] <= ncx;
    if (!dep && !keep) continue;
    if (m < cap) {
      for (int a = 0; a < 3; a++) {
        th[a * cap + m] = a1[a];
        te[a * cap + m] = a2[a];
        v[a * cap + m] = vv[a];
      }
      flag[m] = (uint8_t)(dep | (keep << 1));
    }
    m++;
  }
  return m;
}

void or_set_inflow(orc *h, uint64_t seed, double ppc, double vsw, double vth, double beta_p,
                   uint32_t next_id) {
  h->in_seed = seed;
  h->in_ppc = ppc;
  h->in_vsw = vsw;
  h->in_vth = vth;
  h->in_bp = beta_p;
  h->next_id = next_id;
}

/* append candidate c of the inflow list at cell coordinates t (th or te) */
static void inflow_append(const orc *h, or_state *s, const double *t, const double *v,
                          int64_t cap, int64_t c, uint32_t id) {
  if (s->Ni_tot >= s->Ni_max) return;
  const int M = s->Ni_max, l = s->Ni_tot;
  const real d[3] = {h->dx, h->dy, h->delz};
  const real *q[3] = {h->qx, h->qy, h->qz};
  for (int a = 0; a < 3; a++) {
    s->xp[(size_t)a * M + l] = q[a][0] + (real)t[a * cap + c] * d[a];
    s->vp[(size_t)a * M + l] = (real)v[a * cap + c];
    s->vp1[(size_t)a * M + l] = (real)v[a * cap + c];
  }
  if (s->beta_p) s->beta_p[l] = (real)h->in_bp;
  if (s->mrat) s->mrat[l] = R(1.0);
  if (s->id) s->id[l] = id;
  s->Ni_tot = l + 1;
  s->input_k += or_ke(h, s, l + 1, l + 2);
}

/* kinetic energy without the ion mass: sum 1/2 |v km_to_m|^2 / (mrat beta
   beta_p) over ions l0 <= l < l1 (part_init.f:66-71 with the weight
   1/(beta beta_p) throughout) */
double or_ke(const orc *h, const or_state *s, int l0, int l1) {
  double e = 0;
  const int M = s->Ni_max;
  for (int l = l0; l < l1; l++) {
    double bp = s->beta_p ? (double)s->beta_p[l - 1] : 1.0;
    double mr = s->mrat ? (double)s->mrat[l - 1] : 1.0;
    double v2 = 0;
    for (int m = 1; m <= 3; m++) {
      double v = (double)s->vp[(size_t)(m - 1) * M + (l - 1)] * 1e3;
      v2 += v * v;
    }
    e += 0.5 * v2 / (mr * (double)h->beta * bp);
  }
  return e;
}

/* Energy_diag, part_init.f:14-139: the grid loop of :55-72 over the unique
   nodes, vol = dx_cell dy_cell dz_cell km_to_m^3, m_q = mion/q; constants of
   inputs.f:19,50-57 */
void or_energy_diag(const orc *h, const or_state *s, double mion, const double b0[3],
                    double out[10]) {
  const double q = 1.6e-19, pi = 3.14159, mu0 = pi * 4.0e-7, eps = 8.85e-12, km = 1e3;
  const double m_q = mion / q;
  double Euf = 0, EB1 = 0, EBc[3] = {0, 0, 0}, EE = 0;
  for (int k = 2; k <= h->nz - 1; k++)
    for (int j = 2; j <= h->ny - 1; j++)
      for (int i = 2; i <= h->nx - 1; i++) {
        double vol = (double)h->dx_cell[i - 1] * (double)h->dy_cell[j - 1] *
                     (double)h->dz_cell[k - 1] * km * km * km;
        double denf = (double)s->np[F3(i, j, k)] / (km * km * km);
        for (int m = 1; m <= 3; m++) {
          double b = (double)s->b1[F4(i, j, k, m)];
          double e = (double)s->E[F4(i, j, k, m)];
          double u = (double)s->up[F4(i, j, k, m)];
          EBc[m - 1] += vol / (2.0 * mu0) * (m_q * b) * (m_q * b);
          EB1 += vol / (2.0 * mu0) * (m_q * (b - b0[m - 1])) * (m_q * (b - b0[m - 1]));
          EE += eps * vol / 2.0 * (m_q * e * km) * (m_q * e * km);
          Euf += 0.5 * mion * denf * vol * (u * km) * (u * km);
        }
      }
  double Evp = mion * or_ke(h, s, 1, s->Ni_tot + 1);
  double input = mion * s->input_k, total = Evp + EE + EB1;
  double o[10] = {input, Evp, Euf, EB1, EBc[0], EBc[1], EBc[2], EE, total,
                  input != 0 ? total / input : 0.0};
  memcpy(out, o, sizeof(o));
}

/* the load file of swap_amd/io.py (header "SWAPLD01", n, nc, flags, seed,
   d; cell, frac, vp, vp1 [, beta_p, mrat] [, id]), read with plain fread:
   with cap = 0 only *n and nc are filled; else up to cap ions, arrays as
   [3][cap] (beta_p, mrat, id may be NULL).  Returns 0, or -1 if the file
   is not a load file or is short. */
int or_read_load(const char *path, int64_t cap, int64_t *n, int nc[3], uint32_t *cell,
                 double *frac, double *vp, double *vp1, float *bp, float *mr, uint32_t *id) {
  FILE *f = fopen(path, "rb");
  if (!f) return -1;
  char magic[8];
  int32_t hdr[4];
  int64_t nn;
  uint64_t seed;
  double d[3];
  int ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, "SWAPLD01", 8) == 0 &&
           fread(&nn, 8, 1, f) == 1 && fread(hdr, 4, 4, f) == 4 && fread(&seed, 8, 1, f) == 1 &&
           fread(d, 8, 3, f) == 3;
  if (!ok) { fclose(f); return -1; }
  *n = nn;
  nc[0] = hdr[0]; nc[1] = hdr[1]; nc[2] = hdr[2];
  if (cap == 0) { fclose(f); return 0; }
  int64_t m = nn < cap ? nn : cap;
  ok = fread(cell, 4, (size_t)m, f) == (size_t)m;
  if (ok && m < nn) ok = fseek(f, 4 * (nn - m), SEEK_CUR) == 0;
  double *dst[3] = {frac, vp, vp1};
  for (int q = 0; q < 3 && ok; q++)
    for (int a = 0; a < 3 && ok; a++) {
      ok = fread(dst[q] + (size_t)a * cap, 8, (size_t)m, f) == (size_t)m;
      if (ok && m < nn) ok = fseek(f, 8 * (nn - m), SEEK_CUR) == 0;
    }
  if (ok && (hdr[3] & 1)) {
    float *fd[2] = {bp, mr};
    for (int q = 0; q < 2 && ok; q++) {
      if (fd[q]) ok = fread(fd[q], 4, (size_t)m, f) == (size_t)m;
      else ok = fseek(f, 4 * m, SEEK_CUR) == 0;
      if (ok && m < nn) ok = fseek(f, 4 * (nn - m), SEEK_CUR) == 0;
    }
  }
  if (ok && (hdr[3] & 2) && id) ok = fread(id, 4, (size_t)m, f) == (size_t)m;
  fclose(f);
  return ok ? 0 : -1;
}
