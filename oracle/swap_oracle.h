/*
 * swap_oracle.h -- CPU restatement of SWAP's hybrid-PIC time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP library
 * in swap_amd/csrc.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product path never calls it.
 *
 * One source, two builds: -DOR_REAL=double (liboracle_f64.so, FP64
 * reference arithmetic) and -DOR_REAL=float (liboracle_f32.so, the reference's own
 * arithmetic width, makefile:1 "-real-size 32").
 *
 * Array conventions follow the reference exactly (Fortran column-major):
 *   scalar grid  (nx,ny,nz)      -> C array [nz][ny][nx]
 *   vector grid  (nx,ny,nz,3)    -> C array [3][nz][ny][nx]
 *   particles    (Ni_max,3)      -> C array [3][Ni_max]
 *   weights      (Ni_max,8)      -> C array [8][Ni_max]
 *   ijkp         (Ni_max,3) int  -> C array [3][Ni_max], 1-based values
 * nx,ny,nz include the two ghost layers (reference dimensions.f:5).
 */
#ifndef SWAP_ORACLE_H
#define SWAP_ORACLE_H
#include <stdint.h>

#ifndef OR_REAL
#define OR_REAL double
#endif
typedef OR_REAL real;

typedef struct orc orc;

/* --- context / geometry (initial.f grd7, inputs.f initparameters) --- */
orc *or_create(int nx, int ny, int nz, double dx, double dy, double delz,
               double dt, int ntsub, double alpha, double beta);
void or_destroy(orc *h);
int or_real_size(void);
/* copy the 1-D coordinate arrays out: which = 0..11 (qx,qy,qz,dx_grid,
   dy_grid,dz_grid,dx_cell,dy_cell,dz_cell,xrat,yrat,zrat) */
void or_get_coord(const orc *h, int which, real *out);
void or_set_dtsub(orc *h, double dtsub, double ntf);
void or_get_dtsub(const orc *h, double *dtsub, double *ntf);

/* --- inputs.f constants --- */
void or_initparameters(double b0_init, int ion_amu, double nf_init,
                       double dt_frac, double dx_frac, double nu_init_frac,
                       int ntsub, double *out /* 8: mion,omega_p,lambda_i,
                       dx,dt,dtsub_init,nu_init,alpha */);
void or_grd6_b0(const orc *h, double b0_init, double mion, real *b0);

/* --- boundary.f --- */
void or_periodic(const orc *h, real *b);
void or_periodic_scalar(const orc *h, real *b);

/* --- grid_interp.f --- */
void or_edge_to_center(const orc *h, real *bt, real *btc);
void or_edge_to_face(const orc *h, real *bt, real *btmf);
void or_face_to_center(const orc *h, real *v, real *vc);

/* --- gutsf.f --- */
void or_crossf2(orc *h, real *aa, real *btc, real *cc);
void or_curlB(const orc *h, real *b1, const real *np, real *aj);
void or_curlE(const orc *h, const real *E, real *curl_E);
void or_get_E(orc *h, real *E, real *bt, real *aj, real *up,
              const real *np, const real *nu);
void or_predict_B(orc *h, real *b1, real *b12, real *b1p2, real *bt,
                  real *E, real *aj, real *up, const real *np,
                  const real *nu);
void or_get_Ep1(orc *h, real *E, real *b1, real *b1p2, real *aj,
                real *up, const real *np, const real *nu);
void or_correct_B(orc *h, real *b1, real *b1p2, real *E, real *aj,
                  real *up, const real *np, const real *nu);
void or_f_update_tlev(const orc *h, real *b1, real *b12, const real *b1p2,
                      real *bt);
void or_check_time_step(orc *h, const real *bt, const real *np);

/* --- gutsp.f (particles) --- */
void or_get_interp_weights(const orc *h, int Ni_tot, int Ni_max,
                           const real *xp, int32_t *ijkp, real *wght);
void or_update_np(orc *h, int Ni_tot, int Ni_max, const int32_t *ijkp,
                  const real *wght, const real *beta_p, real *np);
void or_update_up(orc *h, int Ni_tot, int Ni_max, const int32_t *ijkp,
                  const real *wght, const real *beta_p, const real *vp,
                  real *np, real *up);
void or_extrapol_up(orc *h, int Ni_tot, int Ni_max, const int32_t *ijkp,
                    const real *wght, const real *beta_p, const real *vp,
                    const real *vp1, real *np, real *up);
void or_get_Ep(orc *h, int Ni_tot, int Ni_max, const int32_t *ijkp,
               const real *wght, const real *mrat, real *aj, real *up,
               const real *btc, real *Ep);
void or_get_vplus_vminus(const orc *h, int Ni_tot, int Ni_max,
                         const int32_t *ijkp, const real *wght,
                         const real *mrat, const real *Ep, const real *btc,
                         const real *vp, real *vplus, real *vminus);
void or_improve_up(orc *h, int Ni_tot, int Ni_max, const int32_t *ijkp,
                   const real *wght, const real *beta_p, real *vp1,
                   const real *vplus, const real *vminus, real *np,
                   real *up);
void or_get_vp_final(const orc *h, int Ni_tot, int Ni_max, const real *Ep,
                     real *vp, real *vp1, const real *vplus);
void or_move_ion_half(const orc *h, int Ni_tot, int Ni_max, real *xp,
                      const real *vp);

/* --- one full time step in maind.f:385-479 order (no ionization) --- */
typedef struct {
  real *b1, *b12, *b1p2, *bt, *btc, *np, *up, *aj, *E, *nu;
  real *xp, *vp, *vp1, *vplus, *vminus, *Ep, *wght, *beta_p, *mrat;
  int32_t *ijkp;
  uint32_t *id;    /* optional ion ids, carried through or_remove_escaped */
  int Ni_tot, Ni_max;
  int escaped;     /* ions or_step removed through an open x */
  int injected;    /* ions its inflow appended */
  double input_k;  /* input-energy account without the ion mass: + for ions
                      appended, - for ions removed (or_energy_diag) */
} or_state;
void or_step(orc *h, or_state *s);
/* --- open x (build-defined inflow/outflow; the reference is periodic):
       zero-gradient x ghosts, the x ghost layer folded onto layer 2, face
       layers divided by their share of the particle domain, no x wrap in
       move_ion_half, ions outside (q(1), q(nx-1)] removed after each half
       move --- */
/* Energy_diag, part_init.f:14-139, over the unique nodes 2..n-1 and with
   the weight 1/(beta beta_p): out = {input_E, Evp, Euf, EB1, EB1x, EB1y,
   EB1z, EE, total_E, norm_E} in J for ions of mass mion (mion/mrat), with
   input_E = mion * s->input_k; b0: the uniform background field */
void or_energy_diag(const orc *h, const or_state *s, double mion, const double b0[3],
                    double out[10]);
/* kinetic energy / mion of ions [l0, l1) (1-based l0, exclusive l1) */
double or_ke(const orc *h, const or_state *s, int l0, int l1);
void or_set_open_x(orc *h, int on);
/* OpenMP threads of the particle, grid and deposit loops (default 1: the
   serial restatement).  With more threads each deposits its share of the
   particles into its own grid and the grids are added in thread order:
   deterministic for a given count, different rounding across counts. */
void or_set_threads(orc *h, int n);
int or_remove_escaped(const orc *h, or_state *s);
/* inflow through the open x face in every later or_step (ppc = 0: none);
   next_id: the id of the first ion it appends */
void or_set_inflow(orc *h, uint64_t seed, double ppc, double vsw, double vth, double beta_p,
                   uint32_t next_id);

/* --- build-only pieces (no reference counterpart; FP32/integer mirror of
       the device arithmetic so integer bookkeeping can be bit-compared) --- */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                      uint32_t out[4]);
/* split-precision half move: packed cell (i | j<<11 | k<<21), fraction in
   (0,1], s[a] = (dt/2)/d_a as FP32.  periodic in every axis. */
void or_split_move(int64_t n, uint32_t *cell, float *fx, float *fy, float *fz,
                   const float *vx, const float *vy, const float *vz,
                   const float s[3], const int nc[3]);
/* stable counting sort: perm[r] = index of the r-th element in key order */
void or_stable_sort(int64_t n, const uint32_t *key, uint32_t nkeys,
                    uint32_t *perm);

/* seeded drifting-Maxwellian load, split form (cells packed as above,
   fractions and velocities FP64, component-major [3][n]) */
void or_load_maxwellian(int64_t n, uint64_t seed, uint32_t stream, const int nc[3], double vsw,
                        double vth, uint32_t *cell, double *frac, double *v);

/* Ionize_sw_mp (chem_rates.f:508-646) with the build's RNG: returns the
   count dNi and writes the first min(dNi, cap) new ions (cell, frac [3][cap],
   v [3][cap]); see swap_oracle.c */
int64_t or_ionize_sw(const int nc[3], double dx, double dy, double dz, double dt, double beta,
                     double ndot, double vsw, double beta_pu, uint64_t seed, uint32_t step,
                     uint32_t *cell, double *frac, double *v, int64_t cap);

/* pickup-ion cloud (chem_rates.f:290-503, 805-843 restated with the
   build's RNG): cl = {cx, cy, cz, N0, r0, H, tau_photo, rmin, sigma_chex,
   mrat_pu, beta_min}.  Counts and weights per interior node (i fastest), the new
   ions' split positions in node order (m ions, frac [3][m]), and the charge
   exchange decision per existing ion. */
void or_cloud_counts(const int nc[3], double dx, double dy, double dz, double beta, double dt,
                     const double *cl, int max_per_cell, uint64_t seed, uint32_t step,
                     uint32_t *cnt, float *bp);
void or_cloud_ions(const int nc[3], const uint32_t *cnt, uint64_t seed, uint32_t step,
                   int open_x, uint32_t *cell, double *frac, int64_t m);
/* the load file written by swap_amd/io.py (see swap_oracle.c) */
int or_read_load(const char *path, int64_t cap, int64_t *n, int nc[3], uint32_t *cell,
                 double *frac, double *vp, double *vp1, float *bp, float *mr, uint32_t *id);
/* the inflow candidates of one step (build-defined): see swap_oracle.c */
int64_t or_inflow(uint64_t seed, uint32_t step, const int nc[3], double dt, const double d[3],
                  double ppc, double vsw, double vth, double *th, double *te, double *v,
                  uint8_t *flag, int64_t cap);
/* id: the ions' ids (the Philox counter; NULL: the array index) */
void or_chex(int64_t n, const uint32_t *cell, const float *fx, const float *fy, const float *fz,
             const float *vx, const float *vy, const float *vz, const uint32_t *id, double dx,
             double dy, double dz, double dt, const double *cl, uint64_t seed, uint32_t step,
             uint8_t *hit);

#endif
